// Per-precision time-step driver: launches one step's kernels in the
// reference's order (simulation.c:157-222):
//   observers -> Mur E -> E sweep (fast box with the first Debye layer, then
//   the shell: inner-grid edges and UPML) -> remaining Debye layers -> SIBC E
//   -> sources/TF-SF E -> ghost E -> H sweep (fast box, shell: Mur normal H
//   and UPML) -> SIBC H -> sources/TF-SF H -> ghost H
// The aux grids of the plane waves touch nothing but their own arrays: their
// E update rides in the observers' launch and their H update in the H
// corrections' launch, which keeps the order in which each aux array is
// written and read.
#pragma once
#include "vgpu_kernels.h"
#include <algorithm>
#include <map>
#include <string>
#include <vector>

namespace vgpu {

template <typename T> static T *P(void *p) { return static_cast<T *>(p); }

// Sweep arguments of a step that reads field set `in` and writes set
// c->next(in) (the other set with the fused sweep, the same set otherwise).
template <typename T> static void fill_volume_args(Ctx *c, EArgs<T> &ea, HArgs<T> &ha, int in)
{
  ea.NX = ha.NX = c->NX;
  ea.NY = ha.NY = c->NY;
  ea.NZ = ha.NZ = c->NZ;
  ea.PX = ha.PX = c->PX;
  ea.plane = ha.plane = c->plane;
  const int out = c->next(in);
  for (int d = 0; d < 3; ++d) {
    ea.E[d] = P<T>(c->fld(out, d));
    ea.Ein[d] = P<T>(c->fld(in, d));
    ea.H[d] = P<T>(c->fld(in, 3 + d));
    ha.H[d] = P<T>(c->fld(out, 3 + d));
    ha.Hin[d] = P<T>(c->fld(in, 3 + d));
    ha.E[d] = P<T>(c->fld(out, d));
    ea.alpha[d] = P<T>(c->alpha[d]);
    ea.beta[d] = P<T>(c->beta[d]);
    ha.gamma[d] = P<T>(c->gamma[d]);
    ea.idh[d] = P<T>(c->idh[d]);
    ha.ide[d] = P<T>(c->ide[d]);
    for (int f = 0; f < 6; ++f) ea.in[d][f] = c->grid.gfilim[d][f];
    // ordinary H boxes: inner grid plus the normal H of the two Mur faces on this axis
    for (int f = 0; f < 6; ++f) ha.in[d][0][f] = c->grid.gfilim[3 + d][f];
    for (int s = 0; s < 2; ++s) {
      const int face = 2 * d + s;
      if ((c->murmask >> face) & 1u)
        for (int f = 0; f < 6; ++f) ha.in[d][1 + s][f] = c->fmlim[face][3 + d][f];
      else {
        for (int f = 0; f < 6; ++f) ha.in[d][1 + s][f] = 0;
        ha.in[d][1 + s][1] = -1;
      }
    }
    const int pa = 0;
    (void)pa;
    ea.ad[d] = P<T>(c->prof[0][d]);
    ea.bd[d] = P<T>(c->prof[1][d]);
    ea.ah[d] = P<T>(c->prof[2][d]);
    ea.ibh[d] = P<T>(c->prof[5][d]);
    ha.ad[d] = P<T>(c->prof[0][d]);
    ha.ah[d] = P<T>(c->prof[2][d]);
    ha.bh[d] = P<T>(c->prof[3][d]);
    ha.ibd[d] = P<T>(c->prof[4][d]);
  }
  ea.ecls = P<uint16_t>(c->d_ecls); ea.etab = P<T>(c->d_etab); ea.edslot = P<int>(c->d_edslot);
  ha.hcls = P<uint16_t>(c->d_hcls); ha.htab = P<T>(c->d_htab);
  ea.dfuse = c->debyeFused ? 1 : 0;
  if (c->debyeFused) {
    ea.dbase = P<int>(c->d_dbase);
    ea.dlast = P<T>(c->d_dlast); ea.dJ = P<T>(c->d_dJ); ea.dstride = c->debyeSites;
    ea.dalpha = P<double>(c->d_dalpha); ea.dbeta = P<double>(c->d_dbeta); ea.dnp = P<int>(c->d_dnp);
    ea.dmaxP = c->debyeMaxPoles; ea.dcplx = c->debyeCplx;
  }
  ea.pmlmask = ha.pmlmask = c->pmlmask;
  for (int r = 0; r < 6; ++r)
    for (int d = 0; d < 3; ++d) {
      for (int f = 0; f < 6; ++f) {
        ea.pl[r][d][f] = c->fplim[r][d][f];
        ha.pl[r][d][f] = c->fplim[r][3 + d][f];
      }
      ea.PP[r][d] = P<T>(c->aux1[r][d]);
      ea.P[r][d] = P<T>(c->aux2[r][d]);
      ha.B[r][d] = P<T>(c->auxB[r][d]);
    }
  ea.PPpool = P<T>(c->auxPool[0]);
  ea.Ppool = P<T>(c->auxPool[1]);
  ha.Bpool = P<T>(c->auxPool[2]);
}

// UPML auxiliaries of slab r, component C, carried by a shell run.
template <typename T> static void run_aux(const EArgs<T> &a, int r, int C, void *&x, void *&y)
{
  x = a.PP[r][C];
  y = a.P[r][C];
}
template <typename T> static void run_aux(const HArgs<T> &a, int r, int C, void *&x, void *&y)
{
  x = a.B[r][C];
  y = nullptr;
}

// Shell of a sweep: the cells outside its fast box, cut into pieces.  The
// cut planes on y and z are the limits of the fast box, of every component's
// ordinary-update boxes and of its UPML slabs, so over each product of a y
// and a z interval a cell's regions (classify) depend on x alone.  Such a
// product is one row segment in x (two, either side of the fast box, when
// the rows pass through it), whose regions along x form each component's
// runs; products with equal segments and runs are merged along y, then z.
// The launch blocks are then ordered so that block b runs on XCD b % 8 and
// each XCD gets one contiguous eighth of every piece's blocks.
template <int NB, typename A>
static int build_shell(const A &a, const int (*in)[NB][6], const int *fb, const int N[3], std::vector<ShellPiece> &pcs,
                       std::vector<ShellWork> &work, long long &cells, std::string &err, int lw = 0, int kcPlan = 0)
{
  const bool hasFast = fb[0] <= fb[1] && fb[2] <= fb[3] && fb[4] <= fb[5];
  std::vector<int> cut[3];
  for (int ax = 1; ax < 3; ++ax) {
    auto add = [&](int lo, int hi) {
      if (lo > hi) return;
      cut[ax].push_back(std::max(0, std::min(lo, N[ax])));
      cut[ax].push_back(std::max(0, std::min(hi + 1, N[ax])));
    };
    add(0, N[ax] - 1);
    if (hasFast) add(fb[2 * ax], fb[2 * ax + 1]);
    for (int C = 0; C < 3; ++C) {
      for (int b = 0; b < NB; ++b) add(in[C][b][2 * ax], in[C][b][2 * ax + 1]);
      for (int r = 0; r < 6; ++r)
        if ((a.pmlmask >> r) & 1u) add(a.pl[r][C][2 * ax], a.pl[r][C][2 * ax + 1]);
    }
    std::sort(cut[ax].begin(), cut[ax].end());
    cut[ax].erase(std::unique(cut[ax].begin(), cut[ax].end()), cut[ax].end());
  }
  const int ny = (int)cut[1].size() - 1, nz = (int)cut[2].size() - 1;
  std::vector<ShellPiece> raw;
  long long fastCells = 0, idle = 0;
  std::vector<int> reg[3];
  std::vector<long long> pidx[3];
  for (int tz = 0; tz < nz; ++tz)
    for (int ty = 0; ty < ny; ++ty) {
      const int y0 = cut[1][ty], y1 = cut[1][ty + 1] - 1, z0 = cut[2][tz], z1 = cut[2][tz + 1] - 1;
      const long long rows = (long long)(y1 - y0 + 1) * (z1 - z0 + 1);
      // regions along the row, checked at both corners of the product
      for (int C = 0; C < 3; ++C) {
        reg[C].assign(N[0], -1);
        pidx[C].assign(N[0], 0);
        for (int i = 0; i < N[0]; ++i) {
          long long p1 = 0;
          reg[C][i] = classify<NB>(a, in[C], C, i, y0, z0, pidx[C][i]);
          if (classify<NB>(a, in[C], C, i, y1, z1, p1) != reg[C][i]) {
            err = "shell region not constant over a y-z product";
            return -1;
          }
        }
      }
      const bool through = hasFast && y0 >= fb[2] && y1 <= fb[3] && z0 >= fb[4] && z1 <= fb[5];
      int seg[2][2] = {{0, N[0] - 1}, {1, 0}};
      int nseg = 1;
      if (through) {
        seg[0][1] = fb[0] - 1;
        seg[1][0] = fb[1] + 1;
        seg[1][1] = N[0] - 1;
        nseg = 2;
        fastCells += (long long)(fb[1] - fb[0] + 1) * rows;
      }
      for (int sg = 0; sg < nseg; ++sg) {
        int x0 = seg[sg][0], x1 = seg[sg][1];
        if (x0 > x1) continue;
        const long long segCells = (long long)(x1 - x0 + 1) * rows;
        // trim cells with no component updated from both ends
        auto idleAt = [&](int i) { return reg[0][i] < 0 && reg[1][i] < 0 && reg[2][i] < 0; };
        while (x0 <= x1 && idleAt(x0)) ++x0;
        while (x1 >= x0 && idleAt(x1)) --x1;
        long long used = 0;
        // pieces of at most SHELL_RUNS runs per component
        int xs = x0;
        while (xs <= x1) {
          ShellPiece pc{};
          int xe = xs;
          for (; xe <= x1; ++xe) {
            bool fits = true;
            for (int C = 0; C < 3 && fits; ++C) {
              const bool opens = reg[C][xe] >= 0 && (xe == xs || reg[C][xe] != reg[C][xe - 1]);
              if (opens && pc.nrun[C] == SHELL_RUNS) fits = false;
            }
            if (!fits) break;
            for (int C = 0; C < 3; ++C) {
              const int rg = reg[C][xe];
              if (rg < 0) continue;
              if (xe > xs && rg == reg[C][xe - 1]) {
                pc.run[C][pc.nrun[C] - 1].xb = xe;
                continue;
              }
              ShellRun &rn = pc.run[C][pc.nrun[C]++];
              rn.xa = rn.xb = xe;
              rn.reg = rg;
              if (rg < 6) {
                const int *sl = a.pl[rg][C];
                rn.p0 = pidx[C][xe];
                rn.psy = sl[1] - sl[0] + 1;
                rn.psz = rn.psy * (sl[3] - sl[2] + 1);
                run_aux(a, rg, C, rn.a, rn.b);
              }
            }
          }
          pc.box[0] = xs; pc.box[1] = xe - 1;
          pc.box[2] = y0; pc.box[3] = y1;
          pc.box[4] = z0; pc.box[5] = z1;
          raw.push_back(pc);
          used += (long long)(xe - xs) * rows;
          xs = xe;
        }
        idle += segCells - used;
      }
    }
  // merge along y, then z: a piece extends the last one with the same x
  // range and runs and the same extent on the third axis when adjacent
  auto key_of = [](const ShellPiece &pc, int ax) {
    std::vector<long long> key = {pc.box[0], pc.box[1]};
    const int q = ax == 1 ? 2 : 1;
    key.push_back(pc.box[2 * q]);
    key.push_back(pc.box[2 * q + 1]);
    for (int C = 0; C < 3; ++C) {
      key.push_back(pc.nrun[C]);
      for (int r = 0; r < pc.nrun[C]; ++r) {
        key.push_back(pc.run[C][r].xa);
        key.push_back(pc.run[C][r].xb);
        key.push_back(pc.run[C][r].reg);
      }
    }
    return key;
  };
  auto merge = [&](std::vector<ShellPiece> &v, int ax) {
    std::vector<ShellPiece> out;
    std::map<std::vector<long long>, size_t> last;
    for (const ShellPiece &pc : v) {
      const std::vector<long long> key = key_of(pc, ax);
      auto it = last.find(key);
      if (it != last.end() && out[it->second].box[2 * ax + 1] + 1 == pc.box[2 * ax]) {
        out[it->second].box[2 * ax + 1] = pc.box[2 * ax + 1];
      } else {
        last[key] = out.size();
        out.push_back(pc);
      }
    }
    v.swap(out);
  };
  merge(raw, 1);
  merge(raw, 2);
  // planes per block: up to SHELL_KC, fewer when the shell is too small to
  // give about 512 blocks (two per CU) that way.  A thread re-reads H of the
  // plane below its first plane and redoes its region lookup, so one plane
  // per thread costs more than the occupancy it buys: on 1-2 M-cell grids
  // with an 8-cell UPML, two to four planes ran the step 10-14 % faster
  // than the one a 2048-block target gave them
  long long rawCells = 0;
  for (const ShellPiece &pc : raw)
    rawCells += (long long)(pc.box[1] - pc.box[0] + 1) * (pc.box[3] - pc.box[2] + 1) * (pc.box[5] - pc.box[4] + 1);
  int kcmax = (int)std::max(1LL, std::min((long long)SHELL_KC, rawCells / (256LL * 512)));
  if (const char *ev = getenv("VGPU_SHELL_KC")) kcmax = std::max(1, std::min(atoi(ev), 64));  // tuning override
  if (kcPlan > 0) kcmax = kcPlan;  // a plan for launches shared with other grids (group runs)
  cells = 0;
  pcs.clear();
  for (ShellPiece pc : raw) {
    pc.ni = pc.box[1] - pc.box[0] + 1;
    pc.nj = pc.box[3] - pc.box[2] + 1;
    pc.nk = pc.box[5] - pc.box[4] + 1;
    const long long n = (long long)pc.ni * pc.nj * pc.nk;
    if (n > (1LL << 30)) {
      err = "shell piece too large";
      return -1;
    }
    pc.cells = (int)n;
    // Rows shorter than half a line are padded to whole lines of lw
    // elements, each starting on a line, so a wavefront's loads touch one
    // line per row instead of one line per lane (a piece of one- or
    // two-cell rows, as at a Mur face's x end).  Longer rows stay unpadded:
    // there the idle lanes cost more than the split lines (measured).
    const bool pad = lw > 0 && 2 * pc.ni < lw;
    pc.xo = pad ? pc.box[0] % lw : 0;
    pc.nip = pad ? (pc.xo + pc.ni + lw - 1) / lw * lw : pc.ni;
    if ((long long)pc.nip * pc.nj > (1LL << 30)) {
      err = "shell piece too large";
      return -1;
    }
    pc.nxy = pc.nip * pc.nj;
    pc.nxyb = (pc.nxy + 255) / 256;
    pc.kc = std::min(pc.nk, kcmax);
    pc.nblk = pc.nxyb * ((pc.nk + pc.kc - 1) / pc.kc);
    cells += n;
    pcs.push_back(pc);
  }
  if (fastCells + idle + cells != (long long)N[0] * N[1] * N[2]) {
    err = "fast box and shell do not tile the grid";
    return -1;
  }
  // launch order: XCD x runs blocks 8*slot + x, slot over its eighths
  std::vector<ShellWork> per[8];
  for (int q = 0; q < (int)pcs.size(); ++q) {
    const int nb = pcs[q].nblk;
    for (int x = 0; x < 8; ++x)
      for (int b = (int)((long long)nb * x / 8); b < (int)((long long)nb * (x + 1) / 8); ++b) per[x].push_back({q, b});
  }
  size_t len = 0;
  for (int x = 0; x < 8; ++x) len = std::max(len, per[x].size());
  work.assign(8 * len, ShellWork{-1, 0});
  for (int x = 0; x < 8; ++x)
    for (size_t t = 0; t < per[x].size(); ++t) work[8 * t + x] = per[x][t];
  return 0;
}

// The E and H shells of a context (build_shell); kcPlan > 0 fixes the planes
// per block.
template <typename T>
static int plan_shells(Ctx *c, int kcPlan, std::vector<ShellPiece> &ep, std::vector<ShellWork> &ew,
                       std::vector<ShellPiece> &hp, std::vector<ShellWork> &hw, long long cells[2], std::string &err)
{
  EArgs<T> ea{};
  HArgs<T> ha{};
  fill_volume_args<T>(c, ea, ha, 0);
  const int N[3] = {c->NX, c->NY, c->NZ};
  int ein[3][1][6];
  for (int C = 0; C < 3; ++C)
    for (int q = 0; q < 6; ++q) ein[C][0][q] = ea.in[C][q];
  const int lw = getenv("VGPU_SHELL_ROWS") && getenv("VGPU_SHELL_ROWS")[0] == '0' ? 0 : (int)(128 / sizeof(T));
  return build_shell<1>(ea, ein, c->fastE, N, ep, ew, cells[0], err, lw, kcPlan) ||
         build_shell<3>(ha, ha.in, c->fastH, N, hp, hw, cells[1], err, lw, kcPlan);
}

// Once per context (at finalize): the sweeps' argument blocks go to device
// memory, where the shell kernels and the fused Debye step read them.
template <typename T> int prepare_sweeps(Ctx *c)
{
  for (int e = 0; e < 2; ++e) {
    const int *fb = e == 0 ? c->fastE : c->fastH;
    if (fb[0] <= fb[1] && fb[2] <= fb[3] && fb[4] <= fb[5] &&
        (fb[0] < 1 || fb[2] < 1 || fb[4] < 1 || fb[1] > c->NX - 2 || fb[3] > c->NY - 2 || fb[5] > c->NZ - 2))
      return set_error(c, VGPU_EINVAL, "fast box not inside the grid");
  }
  EArgs<T> ea{}, ea1{};
  HArgs<T> ha{}, ha1{};
  fill_volume_args<T>(c, ea, ha, 0);
  fill_volume_args<T>(c, ea1, ha1, 1);
  std::vector<ShellPiece> ep, hp;
  std::vector<ShellWork> ew, hw;
  std::string err;
  if (plan_shells<T>(c, 0, ep, ew, hp, hw, c->shellCells, err)) return set_error(c, VGPU_EINVAL, "%s", err.c_str());
  c->nEpcs = (int)ep.size(); c->nHpcs = (int)hp.size();
  c->nEwork = (int)ew.size(); c->nHwork = (int)hw.size();
  for (int q = 0; q < 2; ++q) {
    c->d_eargs[q] = dev_alloc(c, sizeof(ea));
    c->d_hargs[q] = dev_alloc(c, sizeof(ha));
  }
  c->d_epcs = dev_alloc(c, ep.size() * sizeof(ShellPiece));
  c->d_hpcs = dev_alloc(c, hp.size() * sizeof(ShellPiece));
  c->d_ework = dev_alloc(c, ew.size() * sizeof(ShellWork));
  c->d_hwork = dev_alloc(c, hw.size() * sizeof(ShellWork));
  if (!c->d_eargs[0] || !c->d_hargs[0] || !c->d_eargs[1] || !c->d_hargs[1] || !c->d_epcs || !c->d_hpcs ||
      !c->d_ework || !c->d_hwork)
    return set_error(c, VGPU_ENOMEM, "sweep arguments");
  auto up = [](void *d, const void *h, size_t n) { return n == 0 || hipMemcpy(d, h, n, hipMemcpyHostToDevice) == hipSuccess; };
  if (!up(c->d_eargs[0], &ea, sizeof(ea)) || !up(c->d_hargs[0], &ha, sizeof(ha)) ||
      !up(c->d_eargs[1], &ea1, sizeof(ea1)) || !up(c->d_hargs[1], &ha1, sizeof(ha1)) ||
      !up(c->d_epcs, ep.data(), ep.size() * sizeof(ShellPiece)) ||
      !up(c->d_hpcs, hp.data(), hp.size() * sizeof(ShellPiece)) ||
      !up(c->d_ework, ew.data(), ew.size() * sizeof(ShellWork)) ||
      !up(c->d_hwork, hw.data(), hw.size() * sizeof(ShellWork)))
    return set_error(c, VGPU_EDEVICE, "sweep arguments upload");
  // Debye sites, correction rows and SIBC H rows: each one's coefficients, gathered once
  if (c->debyeSites) {
    const T *b[3] = {P<T>(c->beta[0]), P<T>(c->beta[1]), P<T>(c->beta[2])};
    void *db = dev_alloc(c, sizeof(b));
    if (!db) return set_error(c, VGPU_ENOMEM, "Debye coefficients");
    if (!up(db, b, sizeof(b))) return set_error(c, VGPU_EDEVICE, "Debye coefficients upload");
    const long long n = c->debyeSites;
    hipLaunchKernelGGL((k_debye_gather<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n,
                       P<long long>(c->d_dkey), P<const T *const>(db), c->indexed ? P<uint16_t>(c->d_ecls) : nullptr,
                       c->indexed ? P<T>(c->d_etab) : nullptr, P<T>(c->d_dsbeta));
    if (hipStreamSynchronize(c->stream) != hipSuccess)
      return set_error(c, VGPU_EDEVICE, "Debye coefficients: %s", hipGetErrorString(hipGetLastError()));
  }
  if (c->sibcHRows) {
    const T *g[3] = {P<T>(c->gamma[0]), P<T>(c->gamma[1]), P<T>(c->gamma[2])};
    void *dg = dev_alloc(c, sizeof(g));
    if (!dg) return set_error(c, VGPU_ENOMEM, "SIBC coefficients");
    if (!up(dg, g, sizeof(g))) return set_error(c, VGPU_EDEVICE, "SIBC coefficients upload");
    const long long n = c->sibcHRows;
    hipLaunchKernelGGL((k_corr_gather<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n,
                       P<long long>(c->d_shkey), 3, P<const T *const>(dg), nullptr, P<T>(c->d_shg), nullptr);
    if (hipStreamSynchronize(c->stream) != hipSuccess)
      return set_error(c, VGPU_EDEVICE, "SIBC coefficients: %s", hipGetErrorString(hipGetLastError()));
  }
  for (int e = 0; e < 2; ++e) {
    if (!c->corrRows[e]) continue;
    const T *coef[3], *alpha[3];
    for (int d = 0; d < 3; ++d) {
      coef[d] = P<T>(e == 0 ? c->beta[d] : c->gamma[d]);
      alpha[d] = P<T>(c->alpha[d]);
    }
    void *dc = dev_alloc(c, sizeof(coef)), *da = dev_alloc(c, sizeof(alpha));
    if (!dc || !da) return set_error(c, VGPU_ENOMEM, "correction coefficients");
    if (!up(dc, coef, sizeof(coef)) || !up(da, alpha, sizeof(alpha)))
      return set_error(c, VGPU_EDEVICE, "correction coefficients upload");
    const long long n = c->corrRows[e];
    hipLaunchKernelGGL((k_corr_gather<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n,
                       P<long long>(c->d_ckey[e]), 2, P<const T *const>(dc), e == 0 ? P<const T *const>(da) : nullptr,
                       P<T>(c->d_ccoef[e]), P<T>(c->d_calpha[e]));
    if (hipStreamSynchronize(c->stream) != hipSuccess)
      return set_error(c, VGPU_EDEVICE, "correction coefficients: %s", hipGetErrorString(hipGetLastError()));
  }
  return VGPU_OK;
}

// Argument blocks of one solve's step launches, for both field sets (index:
// the set a step reads), built on the host before a run.
template <typename T> struct StepArgs {
  EArgs<T> ea[2];
  HArgs<T> ha[2];
  dim3 vblk, egrid, hgrid, eg1, hg1;   // fast kernels: block, tile grid, padded 1-D launch
  FastE<T> fe[2];
  FastH<T> fh[2];
  const EArgs<T> *eap[2];              // argument blocks in device memory (shells)
  const HArgs<T> *hap[2];
  const ShellPiece *epc, *hpc;
  const ShellWork *ewk, *hwk;
  LeapArgs<T> la[2];
  FixArgs<T> fx[2];
  dim3 lgrid, lblk;
  MurArgs<T> ma[2];
  dim3 murGrid;
  DebyeArgs<T> da[2];
  SibcEArgs<T> se[2];
  SibcHArgs<T> sh[2];
  CorrArgs<T> ce[2][2];                // [set][E, H]
  AuxArgs<T> xa;
  bool anyAux;
  ObsArgs<T> oa[2];
  int maxnf;
  VolObsArgs<T> va[2];
  unsigned long long *stepp;           // device step counter
};

template <typename T> void fill_step_args(Ctx *c, StepArgs<T> &A, const VisitRun *chk)
{
  const double dt = c->grid.dt;
  // argument sets by the field set a step reads (index 0/1)
  auto &ea = A.ea;
  auto &ha = A.ha;
  for (int q = 0; q < 2; ++q) fill_volume_args<T>(c, ea[q], ha[q], q);
  // Fast kernels: 64x4 tiles (64 x ty) over the fast boxes, marching z in chunks of 4
  // planes (measured best from 272^3 to 386x386x258; 16k blocks and more at those sizes).
  // rows per block (tuning override VGPU_FAST_TY: 1..8)
  int ty = 4;
  if (const char *ev = getenv("VGPU_FAST_TY")) ty = std::max(1, std::min(atoi(ev), 8));
  A.vblk = dim3(64, ty, 1);
  auto fast_grid = [&](const int *fb, int &kchunk) {
    if (fb[0] > fb[1] || fb[2] > fb[3] || fb[4] > fb[5]) return dim3(0, 0, 0);
    const int gx = (fb[1] - fb[0] + 64) / 64, gy = (fb[3] - fb[2] + ty) / ty, nz = fb[5] - fb[4] + 1;
    kchunk = std::min(4, nz);
    if (const char *ev = getenv("VGPU_FAST_KCHUNK"))  // tuning override: planes per block
      kchunk = std::max(1, std::min(atoi(ev), nz));
    return dim3(gx, gy, (nz + kchunk - 1) / kchunk);
  };
  auto &fe = A.fe;
  auto &fh = A.fh;
  int ekc = 1, hkc = 1;
  const dim3 egrid = A.egrid = fast_grid(c->fastE, ekc), hgrid = A.hgrid = fast_grid(c->fastH, hkc);
  // 1-D launches padded to a multiple of eight blocks (XCD order, xcd_block)
  auto pad8 = [](unsigned long long n) { return dim3((unsigned)((n + 7) / 8 * 8)); };
  A.eg1 = pad8((unsigned long long)egrid.x * egrid.y * egrid.z);
  A.hg1 = pad8((unsigned long long)hgrid.x * hgrid.y * hgrid.z);
  for (int q = 0; q < 2; ++q) {
    FastE<T> &e = fe[q];
    FastH<T> &h = fh[q];
    e.NX = h.NX = c->NX; e.NY = h.NY = c->NY; e.NZ = h.NZ = c->NZ; e.PX = h.PX = c->PX;
    e.plane = h.plane = c->plane;
    e.kchunk = ekc; h.kchunk = hkc;
    for (int f = 0; f < 6; ++f) { e.box[f] = c->fastE[f]; h.box[f] = c->fastH[f]; }
    e.nchx = egrid.x; h.nchx = hgrid.x;
    e.gy = egrid.y; e.gz = egrid.z; h.gy = hgrid.y; h.gz = hgrid.z;
    for (int d = 0; d < 3; ++d) {
      e.E[d] = ea[q].E[d]; e.Ein[d] = ea[q].Ein[d]; e.H[d] = ea[q].H[d]; e.idh[d] = ea[q].idh[d];
      e.alpha[d] = ea[q].alpha[d]; e.beta[d] = ea[q].beta[d];
      h.H[d] = ha[q].H[d]; h.Hin[d] = ha[q].Hin[d]; h.E[d] = ha[q].E[d]; h.ide[d] = ha[q].ide[d];
      h.gamma[d] = ha[q].gamma[d];
    }
    e.ecls = ea[q].ecls; e.etab = ea[q].etab; e.dfuse = ea[q].dfuse; e.ap = P<EArgs<T>>(c->d_eargs[q]);
    h.hcls = ha[q].hcls; h.htab = ha[q].htab;
    if (chk)
      for (int d = 0; d < 3; ++d) { e.vis[d] = chk->vis[d]; h.vis[d] = chk->vis[3 + d]; }
  }
  auto &eap = A.eap;
  auto &hap = A.hap;
  for (int q = 0; q < 2; ++q) {
    eap[q] = P<EArgs<T>>(chk ? chk->eargs : c->d_eargs[q]);
    hap[q] = P<HArgs<T>>(chk ? chk->hargs : c->d_hargs[q]);
  }
  A.epc = P<ShellPiece>(c->d_epcs);
  A.hpc = P<ShellPiece>(c->d_hpcs);
  A.ewk = P<ShellWork>(c->d_ework);
  A.hwk = P<ShellWork>(c->d_hwork);
  // fused leapfrog sweep: 4 waves of R rows per block, 63-cell x tiles, kz-plane z chunks
  auto &la = A.la;
  auto &fx = A.fx;
  dim3 &lgrid = A.lgrid;
  lgrid = dim3(0);
  A.lblk = dim3(64, 4, 1);
  if (c->leap) {
    const int *fb = c->fastE;
    if (fb[0] <= fb[1] && fb[2] <= fb[3] && fb[4] <= fb[5]) {
      const int R = c->leapR, kz = std::min(c->leapKZ, fb[5] - fb[4] + 1);
      const int ntx = (fb[1] - fb[0] + 63) / 63, ngy = (fb[3] - fb[2] + 4 * R) / (4 * R), ntz = (fb[5] - fb[4] + kz) / kz;
      lgrid = pad8((unsigned long long)ntx * ngy * ntz);
      for (int q = 0; q < 2; ++q) {
        LeapArgs<T> &l = la[q];
        l.NX = c->NX; l.NY = c->NY; l.NZ = c->NZ; l.PX = c->PX; l.plane = c->plane;
        for (int f = 0; f < 6; ++f) l.box[f] = fb[f];
        l.ntx = ntx; l.ngy = ngy; l.ntz = ntz; l.kz = kz;
        for (int d = 0; d < 3; ++d) {
          l.E[d] = ea[q].E[d]; l.Ein[d] = ea[q].Ein[d]; l.H[d] = ha[q].H[d]; l.Hin[d] = ha[q].Hin[d];
          l.idh[d] = ea[q].idh[d]; l.ide[d] = ha[q].ide[d];
          l.alpha[d] = ea[q].alpha[d]; l.beta[d] = ea[q].beta[d]; l.gamma[d] = ha[q].gamma[d];
        }
        l.ecls = ea[q].ecls; l.etab = ea[q].etab; l.hcls = ha[q].hcls; l.htab = ha[q].htab;
        l.discard = (int)c->ncell;
        if (chk)
          for (int d = 0; d < 6; ++d) l.vis[d] = chk->vis[d];
        FixArgs<T> &x = fx[q];
        x.n = c->nFix; x.ent = P<FixEnt>(c->d_fix);
        x.NY = c->NY; x.PX = c->PX; x.plane = c->plane;
        for (int d = 0; d < 3; ++d) {
          x.H[d] = ha[q].H[d]; x.Hin[d] = ha[q].Hin[d]; x.E[d] = ha[q].E[d]; x.ide[d] = ha[q].ide[d];
          x.gamma[d] = ha[q].gamma[d];
        }
        x.hcls = ha[q].hcls; x.htab = ha[q].htab;
      }
    }
  }
  auto &ma = A.ma;
  int murmax1 = 1, murmax2 = 1;
  if (c->murmask) {
    for (int q = 0; q < 2; ++q) {
      MurArgs<T> &m = ma[q];
      m.NX = c->NX; m.NY = c->NY; m.PX = c->PX; m.plane = c->plane;
      for (int d = 0; d < 3; ++d) {
        m.E[d] = ea[q].E[d]; m.Ein[d] = ea[q].Ein[d]; m.H[d] = ea[q].H[d];
        m.alpha[d] = P<T>(c->alpha[d]); m.beta[d] = P<T>(c->beta[d]); m.idh[d] = P<T>(c->idh[d]);
      }
      if (c->indexed) { m.ecls = ea[q].ecls; m.etab = ea[q].etab; }
      if (chk)
        for (int d = 0; d < 3; ++d) m.vis[d] = chk->vis[d];
    }
    static const int order[6][2] = {{1, 2}, {1, 2}, {0, 2}, {0, 2}, {1, 0}, {1, 0}};
    for (int f = 0; f < 6; ++f) {
      for (int q = 0; q < 2; ++q) ma[q].zeta[f] = T(c->zeta[f]);
      for (int u = 0; u < 2; ++u) {
        const int slot = 2 * f + u;
        if ((c->murmask >> f) & 1u) {
          const int comp = order[f][u];
          for (int q = 0; q < 2; ++q) {
            ma[q].comp[slot] = comp;
            for (int e = 0; e < 6; ++e) ma[q].box[slot][e] = c->fmlim[f][comp][e];
          }
          const int ax = f >> 1, a1 = ax == 0 ? 1 : 0, a2 = ax == 2 ? 1 : 2;
          const int *b = ma[0].box[slot];
          murmax1 = std::max(murmax1, b[2 * a1 + 1] - b[2 * a1] + 1);
          murmax2 = std::max(murmax2, b[2 * a2 + 1] - b[2 * a2] + 1);
        } else {
          for (int q = 0; q < 2; ++q) ma[q].comp[slot] = -1;
        }
      }
    }
  }
  A.murGrid = dim3((murmax1 + 15) / 16, (murmax2 + 15) / 16, 12);

  auto &da = A.da;
  if (c->debyeSites) {
    for (int q = 0; q < 2; ++q) {
      DebyeArgs<T> &a = da[q];
      a.key = P<long long>(c->d_dkey); a.sbeta = P<T>(c->d_dsbeta);
      for (int d = 0; d < 3; ++d) a.E[d] = ea[q].E[d];
      a.lastE = P<T>(c->d_dlast); a.J = P<T>(c->d_dJ); a.stride = c->debyeSites;
      a.dalpha = P<double>(c->d_dalpha); a.dbeta = P<double>(c->d_dbeta); a.npoles = P<int>(c->d_dnp);
      a.maxPoles = c->debyeMaxPoles; a.cplx = c->debyeCplx;
    }
  }

  auto &se = A.se;
  auto &sh = A.sh;
  if (c->sibcFaces) {
    for (int q = 0; q < 2; ++q) {
      SibcEArgs<T> &e = se[q];
      SibcHArgs<T> &h = sh[q];
      e.nfaces = c->sibcFaces; e.ijk = P<int>(c->d_sijk); e.surf = P<int>(c->d_ssurf);
      e.blocks = P<SibcBlock>(c->d_sblocks);
      e.adjA = P<unsigned char>(c->d_sadjA); e.adjB = P<unsigned char>(c->d_sadjB);
      e.normal = P<int>(c->d_snormal); e.A = P<double>(c->d_sA); e.B = P<double>(c->d_sB);
      e.filtset = P<int>(c->d_sset); e.npoles = P<int>(c->d_snp); e.asymp = P<double>(c->d_sasymp);
      e.q = P<double>(c->d_sq); e.maxPoles = c->sibcMaxPoles; e.zeta = P<double>(c->d_szeta);
      e.real = c->sibcReal ? 1 : 0;
      e.factored = c->sibcFactored ? 1 : 0;
      e.fbase = P<int>(c->d_sfbase); e.fratio = P<double>(c->d_sfratio);
      e.old = P<double>(c->d_sold); e.Etan = P<T>(c->d_setan);
      // SIBC E runs after the E sweep: it zeroes the new tangential E and
      // reads the old H (sibc.c:444-586)
      for (int d = 0; d < 3; ++d) { e.E[d] = ea[q].E[d]; e.H[d] = ea[q].H[d]; }
      e.NY = c->NY; e.PX = c->PX; e.plane = c->plane;
      h.nrows = c->sibcHRows; h.rowkey = P<long long>(c->d_shkey);
      h.rowstart = P<int>(c->d_shstart); h.rowg = P<T>(c->d_shg);
      h.ent = P<SibcHEnt>(c->d_shent); h.Etan = P<T>(c->d_setan);
      for (int d = 0; d < 3; ++d) { h.H[d] = ha[q].H[d]; h.ide[d] = P<T>(c->ide[d]); }
    }
  }

  auto &ce = A.ce;
  for (int q = 0; q < 2; ++q)
    for (int e = 0; e < 2; ++e) {
      CorrArgs<T> &a = ce[q][e];
      a = CorrArgs<T>{};
      a.nrows = c->corrRows[e];
      a.rowkey = P<long long>(c->d_ckey[e]);
      a.rowstart = P<int>(c->d_cstart[e]); a.ent = P<CorrEnt>(c->d_cent[e]);
      a.rowcoef = P<T>(c->d_ccoef[e]); a.rowalpha = P<T>(c->d_calpha[e]);
      a.electric = e == 0;
      a.wf = P<WaveformDev>(c->d_wf); a.src = P<SourceDev>(c->d_src); a.pw = P<PwDev>(c->d_pw);
      a.aux = P<T>(c->d_aux); a.auxmax = c->auxmax; a.eta0 = c->grid.eta0;
      // the E list corrects the new E, the H list the new H
      for (int d = 0; d < 3; ++d) a.F[d] = e == 0 ? ea[q].E[d] : ha[q].H[d];
    }

  auto &xa = A.xa;
  bool &anyAux = A.anyAux;
  anyAux = false;
  for (auto &p : c->pw) anyAux |= p.d.use_aux != 0;
  if (anyAux) {
    xa.npw = (int)c->pw.size(); xa.pw = P<PwDev>(c->d_pw); xa.aux = P<T>(c->d_aux);
    xa.pml = P<T>(c->d_auxpml); xa.prof = P<T>(c->d_auxprof); xa.auxmax = c->auxmax;
    xa.wf = P<WaveformDev>(c->d_wf);
  }

  // observers sample the fields a step starts from: the set it reads
  auto &oa = A.oa;
  int &maxnf = A.maxnf;
  maxnf = 1;
  if (!c->pobs.empty()) {
    for (int q = 0; q < 2; ++q) {
      ObsArgs<T> &o = oa[q];
      o.nobs = (int)c->pobs.size(); o.cell = P<long long>(c->d_ocell);
      o.first = P<unsigned long long>(c->d_ofirst); o.last = P<unsigned long long>(c->d_olast);
      o.nfreq = P<int>(c->d_onf); o.dftoff = P<long long>(c->d_ooff); o.omega = P<double>(c->d_oomega);
      o.td = P<double>(c->d_otd); o.tdsteps = c->totalSteps; o.dre = P<double>(c->d_odre);
      o.dim = P<double>(c->d_odim);
      for (int d = 0; d < 6; ++d) o.F[d] = P<T>(c->fld(q, d));
    }
    for (auto &o : c->pobs) maxnf = std::max(maxnf, (int)o.omega.size());
  }
  auto &va = A.va;
  if (c->vnodes) {
    for (int q = 0; q < 2; ++q) {
      VolObsArgs<T> &v = va[q];
      v.nodecell = P<long long>(c->d_vcell); v.nodeobs = P<int>(c->d_vobs); v.nnodes = c->vnodes;
      v.first = P<unsigned long long>(c->d_vfirst); v.last = P<unsigned long long>(c->d_vlast);
      v.stage = P<float>(c->d_vstage);
      v.ring = c->vring;
      for (int d = 0; d < 6; ++d) v.F[d] = P<T>(c->fld(q, d));
    }
  }

  // Time-dependent kernels read the step number from the device counter
  // (step n = counter + offset within this run), so the same launches can be
  // captured once in a graph and replayed for later runs.
  unsigned long long *stepp = A.stepp = P<unsigned long long>(c->d_step);
  for (int q = 0; q < 2; ++q) {
    oa[q].step = stepp; oa[q].dt = dt;
    va[q].step = stepp;
    for (int e = 0; e < 2; ++e) { ce[q][e].step = stepp; ce[q][e].dt = dt; ce[q][e].half = e == 0 ? 0.0 : 0.5; }
  }
  xa.step = stepp; xa.dt = dt;
}

// Ghost planes of one half-step (surface.c:569-812), copied within the field
// set the step writes.
template <typename T> void launch_ghosts(Ctx *c, int electric, int q, hipStream_t s)
{
  for (int f = 0; f < 6; ++f) {
    const int t = c->ghost[f];
    if (!t) continue;
    const int a = f >> 1;
    const int inner = (f & 1) ? c->grid.gobox[f] - 1 : c->grid.gobox[f];
    const int opp = (f & 1) ? c->grid.gobox[f - 1] : c->grid.gobox[f + 1] - 1;
    const int dst = (f & 1) ? c->grid.gobox[f] : c->grid.gobox[f] - 1;
    GhostArgs<T> g{};
    g.NX = c->NX; g.NY = c->NY; g.NZ = c->NZ; g.PX = c->PX; g.plane = c->plane;
    g.axis = a; g.dst = dst;
    for (int d = 0; d < 3; ++d) { g.lo[d] = c->grid.gobox[2 * d] - 1; g.hi[d] = c->grid.gobox[2 * d + 1]; }
    const int a1 = a == 0 ? 1 : 0, a2 = a == 2 ? 1 : 2;
    const dim3 gb(16, 16), gg((g.hi[a1] - g.lo[a1] + 16) / 16, (g.hi[a2] - g.lo[a2] + 16) / 16);
    int comps[2], nc;
    if (electric) { comps[0] = a; nc = 1; }
    else { comps[0] = 3 + (a + 1) % 3; comps[1] = 3 + (a + 2) % 3; nc = 2; }
    // E: PEC copy, PMC negate, periodic copy from the opposite side; H: PMC negate, periodic, PEC copy.
    g.src = t == 3 ? opp : inner;
    g.sign = (t == 2) ? T(-1) : T(1);
    for (int u = 0; u < nc; ++u) {
      g.F = P<T>(c->fld(c->next(q), comps[u]));
      hipLaunchKernelGGL(k_ghost<T>, gg, gb, 0, s, g);
    }
  }
}

// The loop (simulation.c:136-230).  With chk, one step of the field sweeps
// only (Mur E, the E sweep, the H sweep), in their update-counting form.
// With the fused sweep, step q of the run reads field set c->par ^ (q & 1)
// and writes the other, and the caller flips c->par by the parity of nsteps
// once the run is enqueued; the separate sweeps read and write set 0.
template <typename T> int run_steps(Ctx *c, uint64_t nsteps, hipStream_t s, bool timeKernels, const VisitRun *chk)
{
  const double pi = 4.0 * atan(1.0);
  timeKernels = timeKernels && !c->evE.empty();
  StepArgs<T> A{};
  fill_step_args<T>(c, A, chk);
  const auto &fe = A.fe;
  const auto &fh = A.fh;
  const dim3 egrid = A.egrid, hgrid = A.hgrid, eg1 = A.eg1, hg1 = A.hg1, vblk = A.vblk;
  const auto &eap = A.eap;
  const auto &hap = A.hap;
  const ShellPiece *epc = A.epc, *hpc = A.hpc;
  const ShellWork *ewk = A.ewk, *hwk = A.hwk;
  auto fast_e = [&](int q) {
    const FastE<T> &f = fe[q];
    if (egrid.x && chk) {
      if (c->indexed) hipLaunchKernelGGL((k_e_fast<T, true, false, true>), eg1, vblk, 0, s, f);
      else hipLaunchKernelGGL((k_e_fast<T, false, false, true>), eg1, vblk, 0, s, f);
    } else if (egrid.x) {
      if (c->indexed && c->debyeFused) hipLaunchKernelGGL((k_e_fast<T, true, true>), eg1, vblk, 0, s, f);
      else if (c->indexed) hipLaunchKernelGGL((k_e_fast<T, true, false>), eg1, vblk, 0, s, f);
      else hipLaunchKernelGGL((k_e_fast<T, false, false>), eg1, vblk, 0, s, f);
    }
  };
  auto shell_e = [&](int q) {
    if (c->nEwork) {
      const dim3 g((unsigned)c->nEwork);
      const EArgs<T> *a = eap[q];
      if (chk && c->indexed) hipLaunchKernelGGL((k_e_shell<T, true, true>), g, dim3(256), 0, s, epc, ewk, a);
      else if (chk) hipLaunchKernelGGL((k_e_shell<T, false, true>), g, dim3(256), 0, s, epc, ewk, a);
      else if (c->indexed && !c->leap) hipLaunchKernelGGL((k_e_shell<T, true, false, false>), g, dim3(256), 0, s, epc, ewk, a);
      else if (c->indexed) hipLaunchKernelGGL((k_e_shell<T, true>), g, dim3(256), 0, s, epc, ewk, a);
      else if (!c->leap) hipLaunchKernelGGL((k_e_shell<T, false, false, false>), g, dim3(256), 0, s, epc, ewk, a);
      else hipLaunchKernelGGL((k_e_shell<T, false>), g, dim3(256), 0, s, epc, ewk, a);
    }
  };
  auto fast_h = [&](int q) {
    const FastH<T> &f = fh[q];
    if (hgrid.x && chk) {
      if (c->indexed) hipLaunchKernelGGL((k_h_fast<T, true, true>), hg1, vblk, 0, s, f);
      else hipLaunchKernelGGL((k_h_fast<T, false, true>), hg1, vblk, 0, s, f);
    } else if (hgrid.x) {
      if (c->indexed) hipLaunchKernelGGL((k_h_fast<T, true>), hg1, vblk, 0, s, f);
      else hipLaunchKernelGGL((k_h_fast<T, false>), hg1, vblk, 0, s, f);
    }
  };
  auto shell_h = [&](int q) {
    if (c->nHwork) {
      const dim3 g((unsigned)c->nHwork);
      const HArgs<T> *a = hap[q];
      if (chk && c->indexed) hipLaunchKernelGGL((k_h_shell<T, true, true>), g, dim3(256), 0, s, hpc, hwk, a);
      else if (chk) hipLaunchKernelGGL((k_h_shell<T, false, true>), g, dim3(256), 0, s, hpc, hwk, a);
      else if (c->indexed) hipLaunchKernelGGL((k_h_shell<T, true>), g, dim3(256), 0, s, hpc, hwk, a);
      else hipLaunchKernelGGL((k_h_shell<T, false>), g, dim3(256), 0, s, hpc, hwk, a);
    }
  };

  const auto &la = A.la;
  const auto &fx = A.fx;
  const dim3 lgrid = A.lgrid, lblk = A.lblk;
  auto leap = [&](int q) {
    if (!lgrid.x) return;
    const LeapArgs<T> &l = la[q];
    const int R = c->leapR;
    if (chk) {
      if (c->indexed) {
        if (R == 1) hipLaunchKernelGGL((k_leap<T, true, 1, true>), lgrid, lblk, 0, s, l);
        else if (R == 2) hipLaunchKernelGGL((k_leap<T, true, 2, true>), lgrid, lblk, 0, s, l);
        else hipLaunchKernelGGL((k_leap<T, true, 4, true>), lgrid, lblk, 0, s, l);
      } else {
        if (R == 1) hipLaunchKernelGGL((k_leap<T, false, 1, true>), lgrid, lblk, 0, s, l);
        else if (R == 2) hipLaunchKernelGGL((k_leap<T, false, 2, true>), lgrid, lblk, 0, s, l);
        else hipLaunchKernelGGL((k_leap<T, false, 4, true>), lgrid, lblk, 0, s, l);
      }
    } else if (c->indexed) {
      if (R == 1) hipLaunchKernelGGL((k_leap<T, true, 1>), lgrid, lblk, 0, s, l);
      else if (R == 2) hipLaunchKernelGGL((k_leap<T, true, 2>), lgrid, lblk, 0, s, l);
      else hipLaunchKernelGGL((k_leap<T, true, 4>), lgrid, lblk, 0, s, l);
    } else {
      if (R == 1) hipLaunchKernelGGL((k_leap<T, false, 1>), lgrid, lblk, 0, s, l);
      else if (R == 2) hipLaunchKernelGGL((k_leap<T, false, 2>), lgrid, lblk, 0, s, l);
      else hipLaunchKernelGGL((k_leap<T, false, 4>), lgrid, lblk, 0, s, l);
    }
  };
  auto fix = [&](int q) {
    if (!lgrid.x || !c->nFix) return;
    const dim3 g((unsigned)((c->nFix + 3) / 4));
    if (c->indexed) hipLaunchKernelGGL((k_h_fix<T, true>), g, lblk, 0, s, fx[q]);
    else hipLaunchKernelGGL((k_h_fix<T, false>), g, lblk, 0, s, fx[q]);
  };

  const auto &ma = A.ma;
  const dim3 murGrid = A.murGrid;
  if (chk) {
    const int q = c->par;
    if (c->murmask) hipLaunchKernelGGL((k_mur_e<T, true>), murGrid, dim3(16, 16), 0, s, ma[q]);
    if (c->leap) {
      leap(q);
      shell_e(q);
      shell_h(q);
    } else {
      fast_e(q);
      shell_e(q);
      fast_h(q);
      shell_h(q);
    }
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) return set_error(c, VGPU_EDEVICE, "kernel launch failed: %s", hipGetErrorString(err));
    return VGPU_OK;
  }

  auto &da = A.da;
  const auto &se = A.se;
  const auto &sh = A.sh;
  auto &ce = A.ce;
  auto &xa = A.xa;
  const bool anyAux = A.anyAux;
  auto &oa = A.oa;
  const int maxnf = A.maxnf;
  auto &va = A.va;
  unsigned long long *stepp = A.stepp;
  auto ghost = [&](int electric, int q) { launch_ghosts<T>(c, electric, q, s); };
  int ne = 0, nh = 0;
  for (uint64_t n = 0; n < nsteps; ++n) {
    const int soff = (int)n;
    const int q = c->leap ? c->par ^ (int)(n & 1) : c->par;  // field set this step reads
    // observers, and the aux grids' E update as extra blocks (k_head)
    if (oa[q].nobs || anyAux) {
      oa[q].soff = soff;
      xa.soff = soff;
      const unsigned gy = (unsigned)(oa[q].nobs + (anyAux ? xa.npw : 0));
      hipLaunchKernelGGL(k_head<T>, dim3(oa[q].nobs ? (maxnf + 63) / 64 : 1, gy), dim3(64), 0, s, oa[q], xa, pi);
    }
    if (va[q].nnodes) {
      va[q].soff = soff;
      hipLaunchKernelGGL(k_observe_volume<T>, dim3((unsigned)((va[q].nnodes + 255) / 256)), dim3(256), 0, s, va[q]);
    }
    if (c->murmask) hipLaunchKernelGGL(k_mur_e<T>, murGrid, dim3(16, 16), 0, s, ma[q]);
    // events bracket the fast kernels (the dominant launches): the fused
    // sweep and its fix-up, or the fast E and H kernels
    if (timeKernels && ne < (int)c->evE.size() / 2) hipEventRecord(c->evE[2 * ne], s);
    if (c->leap) leap(q);
    else fast_e(q);
    if (timeKernels && ne < (int)c->evE.size() / 2) { hipEventRecord(c->evE[2 * ne + 1], s); ++ne; }
    shell_e(q);
    // Layers 0 and 1 hold each cell's first block, inside and outside the
    // fast E box, so their sites never share an edge: without fusion they
    // go in one launch.  Later layers follow one launch each, in block order.
    for (int l = c->debyeFused ? 1 : 0; l < c->debyeLayers;) {
      const int l1 = (!c->debyeFused && l == 0 && c->debyeLayers > 1) ? 2 : l + 1;
      da[q].base = c->debyeLayerStart[l];
      da[q].n = c->debyeLayerStart[l1] - c->debyeLayerStart[l];
      l = l1;
      if (da[q].n == 0) continue;
      hipLaunchKernelGGL(k_debye<T>, dim3((unsigned)((da[q].n + 255) / 256)), dim3(256), 0, s, da[q]);
    }
    if (c->sibcFaces) hipLaunchKernelGGL(k_sibc_e<T>, dim3((unsigned)c->sibcBlocks), dim3(128), 0, s, se[q]);
    if (ce[q][0].nrows) {
      ce[q][0].soff = soff;
      hipLaunchKernelGGL(k_corr<T>, dim3((unsigned)((ce[q][0].nrows + 255) / 256)), dim3(256), 0, s, ce[q][0], pi);
    }
    ghost(1, q);
    if (c->leap) shell_h(q);
    if (timeKernels && nh < (int)c->evH.size() / 2) hipEventRecord(c->evH[2 * nh], s);
    if (c->leap) fix(q);
    else fast_h(q);
    if (timeKernels && nh < (int)c->evH.size() / 2) { hipEventRecord(c->evH[2 * nh + 1], s); ++nh; }
    if (!c->leap) shell_h(q);
    if (c->sibcFaces)
      hipLaunchKernelGGL(k_sibc_h<T>, dim3((unsigned)((sh[q].nrows + 255) / 256)), dim3(256), 0, s, sh[q]);
    // H corrections, and the aux grids' H update as extra blocks
    if (ce[q][1].nrows || anyAux) {
      ce[q][1].soff = soff;
      const int nb = (int)((ce[q][1].nrows + 255) / 256);
      const dim3 g((unsigned)(nb + (anyAux ? xa.npw : 0)));
      if (anyAux) hipLaunchKernelGGL(k_corr_aux<T>, g, dim3(256), 0, s, ce[q][1], pi, xa, nb);
      else hipLaunchKernelGGL(k_corr<T>, g, dim3(256), 0, s, ce[q][1], pi);
    }
    ghost(0, q);
  }
  hipLaunchKernelGGL(k_advance_step, dim3(1), dim3(64), 0, s, stepp, (unsigned long long)nsteps);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return set_error(c, VGPU_EDEVICE, "kernel launch failed: %s", hipGetErrorString(err));
  return VGPU_OK;
}

// A solve's argument blocks for packed launches (field set 0), uploaded once
// after set-up.  Contexts with the fused sweep take no part in group runs.
template <typename T> int upload_pack(Ctx *c)
{
  StepArgs<T> A{};
  fill_step_args<T>(c, A, nullptr);
  // shells at SHELL_KC planes per block: a packed launch fills the device
  // with many grids' blocks, where a small grid's own plan used one plane
  std::vector<ShellPiece> ep, hp;
  std::vector<ShellWork> ew, hw;
  std::string err;
  long long cells[2];
  if (plan_shells<T>(c, SHELL_KC, ep, ew, hp, hw, cells, err)) return set_error(c, VGPU_EINVAL, "%s", err.c_str());
  c->d_epcsPk = dev_alloc(c, ep.size() * sizeof(ShellPiece));
  c->d_hpcsPk = dev_alloc(c, hp.size() * sizeof(ShellPiece));
  c->d_eworkPk = dev_alloc(c, ew.size() * sizeof(ShellWork));
  c->d_hworkPk = dev_alloc(c, hw.size() * sizeof(ShellWork));
  if (!c->d_epcsPk || !c->d_hpcsPk || !c->d_eworkPk || !c->d_hworkPk)
    return set_error(c, VGPU_ENOMEM, "group-run shells");
  auto up = [](void *d, const void *h, size_t n) {
    return n == 0 || hipMemcpy(d, h, n, hipMemcpyHostToDevice) == hipSuccess;
  };
  if (!up(c->d_epcsPk, ep.data(), ep.size() * sizeof(ShellPiece)) ||
      !up(c->d_hpcsPk, hp.data(), hp.size() * sizeof(ShellPiece)) ||
      !up(c->d_eworkPk, ew.data(), ew.size() * sizeof(ShellWork)) ||
      !up(c->d_hworkPk, hw.data(), hw.size() * sizeof(ShellWork)))
    return set_error(c, VGPU_EDEVICE, "group-run shells upload");
  c->nEworkPk = (int)ew.size();
  c->nHworkPk = (int)hw.size();
  PackArgs<T> p{};
  p.fe = A.fe[0];
  p.fh = A.fh[0];
  p.epc = P<ShellPiece>(c->d_epcsPk); p.hpc = P<ShellPiece>(c->d_hpcsPk);
  p.ewk = P<ShellWork>(c->d_eworkPk); p.hwk = P<ShellWork>(c->d_hworkPk);
  p.ea = A.eap[0]; p.ha = A.hap[0];
  p.da = A.da[0];
  p.ce = A.ce[0][0];
  p.ch = A.ce[0][1];
  p.oa = A.oa[0];
  p.xa = A.xa;
  p.va = A.va[0];
  p.headGx = A.oa[0].nobs ? (A.maxnf + 63) / 64 : 1;
  c->d_pack = dev_alloc(c, sizeof(p));
  if (!c->d_pack) return set_error(c, VGPU_ENOMEM, "group-run arguments");
  if (hipMemcpy(c->d_pack, &p, sizeof(p), hipMemcpyHostToDevice) != hipSuccess)
    return set_error(c, VGPU_EDEVICE, "group-run arguments: %s", hipGetErrorString(hipGetLastError()));
  return VGPU_OK;
}

// Steps of several solves in lockstep on one stream (vgpu_run_group): the
// same kernels in the same order per solve as run_steps, with the fast and
// shell sweeps, Debye layers, corrections and observers of all solves in one
// packed launch each per half-step; Mur, SIBC and ghost launches stay one
// per solve.  Solves share no memory, so only each solve's own order
// matters, and every block does what it does in the solve's own launch.
// Every context uses the separate sweeps (set 0) and has its PackArgs.
template <typename T> int run_group_steps(Ctx *const *cs, int n, uint64_t nsteps, hipStream_t s)
{
  const double pi = 4.0 * atan(1.0);
  std::vector<StepArgs<T>> A(n);
  for (int i = 0; i < n; ++i) fill_step_args<T>(cs[i], A[i], nullptr);
  // One packed launch: the solves with a positive block count, in order.
  auto plan = [&](auto blocks) {
    PackMap m{};
    unsigned tot = 0;
    for (int i = 0; i < n; ++i) {
      long long base = 0, cnt = 0;
      const long long k = blocks(i, base, cnt);
      if (k <= 0) continue;
      m.args[m.n] = cs[i]->d_pack;
      m.base[m.n] = base;
      m.cnt[m.n] = cnt;
      m.start[m.n++] = tot;
      tot += (unsigned)k;
    }
    m.start[m.n] = tot;
    return m;
  };
  PackMap fastE[2], fastH[2], shellE[2], shellH[2];
  for (int x = 0; x < 2; ++x) {
    const bool idx = x == 0;
    fastE[x] = plan([&](int i, long long &, long long &) {
      return cs[i]->indexed == idx && A[i].egrid.x ? (long long)A[i].eg1.x : 0LL; });
    fastH[x] = plan([&](int i, long long &, long long &) {
      return cs[i]->indexed == idx && A[i].hgrid.x ? (long long)A[i].hg1.x : 0LL; });
    shellE[x] = plan([&](int i, long long &, long long &) {
      return cs[i]->indexed == idx ? (long long)cs[i]->nEworkPk : 0LL; });
    shellH[x] = plan([&](int i, long long &, long long &) {
      return cs[i]->indexed == idx ? (long long)cs[i]->nHworkPk : 0LL; });
  }
  unsigned headGx = 1;
  for (int i = 0; i < n; ++i)
    if (A[i].oa[0].nobs) headGx = std::max(headGx, (unsigned)((A[i].maxnf + 63) / 64));
  PackMap head = plan([&](int i, long long &, long long &) {
    return A[i].oa[0].nobs || A[i].anyAux ? (long long)A[i].oa[0].nobs + (A[i].anyAux ? A[i].xa.npw : 0) : 0LL; });
  PackMap vol = plan([&](int i, long long &, long long &) { return (A[i].va[0].nnodes + 255) / 256; });
  PackMap corrE = plan([&](int i, long long &, long long &) { return (A[i].ce[0][0].nrows + 255) / 256; });
  PackMap corrH = plan([&](int i, long long &, long long &cnt) {
    cnt = (A[i].ce[0][1].nrows + 255) / 256;
    return A[i].ce[0][1].nrows || A[i].anyAux ? cnt + (A[i].anyAux ? A[i].xa.npw : 0) : 0LL; });
  // Debye launches of each solve (run_steps' order: layers 0 and 1 together,
  // then one per layer), packed by launch number
  std::vector<std::vector<std::pair<long long, long long>>> dl(n);
  size_t ndl = 0;
  for (int i = 0; i < n; ++i) {
    const Ctx *c = cs[i];
    for (int l = 0; l < c->debyeLayers;) {
      const int l1 = (l == 0 && c->debyeLayers > 1) ? 2 : l + 1;
      const long long b = c->debyeLayerStart[l], k = c->debyeLayerStart[l1] - b;
      l = l1;
      if (k) dl[i].push_back({b, k});
    }
    ndl = std::max(ndl, dl[i].size());
  }
  std::vector<PackMap> debye(ndl);
  for (size_t j = 0; j < ndl; ++j)
    debye[j] = plan([&](int i, long long &base, long long &cnt) {
      if (j >= dl[i].size()) return 0LL;
      base = dl[i][j].first;
      cnt = dl[i][j].second;
      return (cnt + 255) / 256; });
  const dim3 vblk = A[0].vblk;
  auto blocks = [](const PackMap &m) { return dim3(m.start[m.n]); };  // a packed launch's 1-D grid
  for (uint64_t st = 0; st < nsteps; ++st) {
    const int soff = (int)st;
    if (head.n) {
      head.soff = soff;
      hipLaunchKernelGGL(k_head_pk<T>, dim3(headGx, head.start[head.n]), dim3(64), 0, s, head, pi);
    }
    if (vol.n) {
      vol.soff = soff;
      hipLaunchKernelGGL(k_observe_volume_pk<T>, blocks(vol), dim3(256), 0, s, vol);
    }
    for (int i = 0; i < n; ++i)
      if (cs[i]->murmask) hipLaunchKernelGGL(k_mur_e<T>, A[i].murGrid, dim3(16, 16), 0, s, A[i].ma[0]);
    if (fastE[0].n) hipLaunchKernelGGL((k_e_fast_pk<T, true>), blocks(fastE[0]), vblk, 0, s, fastE[0]);
    if (fastE[1].n) hipLaunchKernelGGL((k_e_fast_pk<T, false>), blocks(fastE[1]), vblk, 0, s, fastE[1]);
    if (shellE[0].n) hipLaunchKernelGGL((k_e_shell_pk<T, true>), blocks(shellE[0]), dim3(256), 0, s, shellE[0]);
    if (shellE[1].n) hipLaunchKernelGGL((k_e_shell_pk<T, false>), blocks(shellE[1]), dim3(256), 0, s, shellE[1]);
    for (const PackMap &m : debye) hipLaunchKernelGGL(k_debye_pk<T>, blocks(m), dim3(256), 0, s, m);
    for (int i = 0; i < n; ++i)
      if (cs[i]->sibcFaces)
        hipLaunchKernelGGL(k_sibc_e<T>, dim3((unsigned)cs[i]->sibcBlocks), dim3(128), 0, s, A[i].se[0]);
    if (corrE.n) {
      corrE.soff = soff;
      hipLaunchKernelGGL(k_corr_e_pk<T>, blocks(corrE), dim3(256), 0, s, corrE, pi);
    }
    for (int i = 0; i < n; ++i) launch_ghosts<T>(cs[i], 1, 0, s);
    if (fastH[0].n) hipLaunchKernelGGL((k_h_fast_pk<T, true>), blocks(fastH[0]), vblk, 0, s, fastH[0]);
    if (fastH[1].n) hipLaunchKernelGGL((k_h_fast_pk<T, false>), blocks(fastH[1]), vblk, 0, s, fastH[1]);
    if (shellH[0].n) hipLaunchKernelGGL((k_h_shell_pk<T, true>), blocks(shellH[0]), dim3(256), 0, s, shellH[0]);
    if (shellH[1].n) hipLaunchKernelGGL((k_h_shell_pk<T, false>), blocks(shellH[1]), dim3(256), 0, s, shellH[1]);
    for (int i = 0; i < n; ++i)
      if (cs[i]->sibcFaces)
        hipLaunchKernelGGL(k_sibc_h<T>, dim3((unsigned)((A[i].sh[0].nrows + 255) / 256)), dim3(256), 0, s, A[i].sh[0]);
    if (corrH.n) {
      corrH.soff = soff;
      hipLaunchKernelGGL(k_corr_h_pk<T>, blocks(corrH), dim3(256), 0, s, corrH, pi);
    }
    for (int i = 0; i < n; ++i) launch_ghosts<T>(cs[i], 0, 0, s);
  }
  for (int i = 0; i < n; ++i)
    hipLaunchKernelGGL(k_advance_step, dim3(1), dim3(64), 0, s, A[i].stepp, (unsigned long long)nsteps);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return set_error(cs[0], VGPU_EDEVICE, "kernel launch failed: %s", hipGetErrorString(err));
  return VGPU_OK;
}

// Update-once check (the reference's CHECK_LIMITS build: grid.h:181-191, and
// checkGrid, grid.c:951-1025).  One step of the field sweeps -- Mur E, the E
// sweep, the H sweep, as the loop launches them -- runs in its counting form,
// which adds one to a per-cell count for each update instead of storing it;
// each component's counts are then tallied against its outer field limits
// folim.  Fields, UPML auxiliaries and Debye state are left as they were.
// Bytes of device memory one update-once check uses: visit counts per
// component (device layout), the two sweep argument blocks, the tally and,
// with counts, their staging in the counts layout; each part 256-byte aligned.
template <typename T> size_t check_limits_bytes(const Ctx *c, bool counts, size_t off[4])
{
  auto al = [](size_t n) { return (n + 255) / 256 * 256; };
  const size_t cells = (size_t)c->plane * c->NZ, ncell = (size_t)c->NX * c->NY * c->NZ;
  off[0] = 0;
  off[1] = al(6 * cells * sizeof(unsigned));
  off[2] = off[1] + al(al(sizeof(EArgs<T>)) + sizeof(HArgs<T>));
  off[3] = off[2] + al(24 * sizeof(unsigned long long));
  return off[3] + (counts ? al(6 * ncell) : 0);
}

template <typename T>
int check_limits(Ctx *c, const int folim[6][6], vgpu_limits_report *rep, unsigned char *counts, void *scratch,
                 size_t scratchBytes)
{
  const size_t cells = (size_t)c->plane * c->NZ;      // device layout, padded rows
  const size_t ncell = (size_t)c->NX * c->NY * c->NZ; // counts layout, x fastest
  unsigned *vis = nullptr;
  char *args = nullptr;
  unsigned long long *tally = nullptr;
  unsigned char *dcounts = nullptr;
  int r = VGPU_OK;
  const size_t hoff = (sizeof(EArgs<T>) + 255) / 256 * 256;
  size_t off[4];
  const size_t need = check_limits_bytes<T>(c, counts != nullptr, off);
  if (scratch && need > scratchBytes)
    return set_error(c, VGPU_EINVAL,
                     "update-once check of this grid needs %zu bytes of scratch, more than the %zu the "
                     "no-allocation entry point holds",
                     need, scratchBytes);
  if (hipDeviceSynchronize() != hipSuccess) return set_error(c, VGPU_EDEVICE, "device error before the check");
  bool have;
  if (scratch) {
    // the caller's buffer: no allocation
    char *b = static_cast<char *>(scratch);
    vis = reinterpret_cast<unsigned *>(b + off[0]);
    args = b + off[1];
    tally = reinterpret_cast<unsigned long long *>(b + off[2]);
    dcounts = counts ? reinterpret_cast<unsigned char *>(b + off[3]) : nullptr;
    have = true;
  } else {
    have = hipMalloc((void **)&vis, 6 * cells * sizeof(unsigned)) == hipSuccess &&
           hipMalloc((void **)&args, hoff + sizeof(HArgs<T>)) == hipSuccess &&
           hipMalloc((void **)&tally, 24 * sizeof(unsigned long long)) == hipSuccess &&
           (!counts || hipMalloc((void **)&dcounts, 6 * ncell) == hipSuccess);
  }
  if (!have) {
    r = set_error(c, VGPU_ENOMEM, "update-once check buffers");
  } else {
    EArgs<T> ea;
    HArgs<T> ha;
    unsigned long long init[24];
    for (int q = 0; q < 24; ++q) init[q] = (q & 3) == 3 ? ~0ull : 0ull;
    bool ok = hipMemcpy(&ea, c->d_eargs[c->par], sizeof ea, hipMemcpyDeviceToHost) == hipSuccess &&
              hipMemcpy(&ha, c->d_hargs[c->par], sizeof ha, hipMemcpyDeviceToHost) == hipSuccess;
    VisitRun vr{};
    for (int q = 0; q < 6; ++q) vr.vis[q] = vis + q * cells;
    for (int d = 0; d < 3; ++d) { ea.vis[d] = vr.vis[d]; ha.vis[d] = vr.vis[3 + d]; }
    vr.eargs = args;
    vr.hargs = args + hoff;
    // the counts start at zero (a fill kernel: no memset node if captured)
    const size_t nv = 6 * cells;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<size_t>((nv + 255) / 256, 65535)), dim3(256), 0, c->stream,
                       vis, nv, 0u);
    ok = ok && hipStreamSynchronize(c->stream) == hipSuccess &&
         hipMemcpy(args, &ea, sizeof ea, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(args + hoff, &ha, sizeof ha, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(tally, init, sizeof init, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) r = set_error(c, VGPU_EDEVICE, "update-once check set-up");
    if (!r) r = run_steps<T>(c, 1, c->stream, false, &vr);
    if (!r) {
      for (int C = 0; C < 6; ++C) {
        Lim6 lim;
        for (int f = 0; f < 6; ++f) lim.v[f] = folim[C][f];
        hipLaunchKernelGGL(k_visit_tally, dim3((c->NX + 255) / 256, c->NY, c->NZ), dim3(256), 0, c->stream,
                           vr.vis[C], c->NX, c->NY, c->NZ, (long long)c->PX, lim, tally + 4 * C,
                           dcounts ? dcounts + C * ncell : nullptr);
      }
      unsigned long long out[24];
      if (hipStreamSynchronize(c->stream) != hipSuccess ||
          hipMemcpy(out, tally, sizeof out, hipMemcpyDeviceToHost) != hipSuccess ||
          (counts && hipMemcpy(counts, dcounts, 6 * ncell, hipMemcpyDeviceToHost) != hipSuccess)) {
        r = set_error(c, VGPU_EDEVICE, "update-once check: %s", hipGetErrorString(hipGetLastError()));
      } else {
        for (int C = 0; C < 6; ++C) {
          rep->missed[C] = out[4 * C];
          rep->repeated[C] = out[4 * C + 1];
          rep->outside[C] = out[4 * C + 2];
          rep->first[C] = out[4 * C + 3] == ~0ull ? -1 : (int64_t)out[4 * C + 3];
        }
      }
    }
  }
  if (!scratch) {
    (void)hipFree(vis);
    (void)hipFree(args);
    (void)hipFree(tally);
    (void)hipFree(dcounts);
  }
  return r;
}

}  // namespace vgpu
