// Per-precision time-step driver: launches one step's kernels in the
// reference's order (simulation.c:157-222):
//   observers -> Mur E -> E sweep (fast box with the first Debye layer, then
//   the shell: inner-grid edges and UPML) -> remaining Debye layers -> SIBC E
//   -> aux grids E -> sources/TF-SF E -> ghost E -> H sweep (fast box, shell:
//   Mur normal H and UPML) -> SIBC H -> aux grids H -> sources/TF-SF H ->
//   ghost H
#pragma once
#include "vgpu_kernels.h"

namespace vgpu {

template <typename T> static T *P(void *p) { return static_cast<T *>(p); }

template <typename T> static void fill_volume_args(Ctx *c, EArgs<T> &ea, HArgs<T> &ha)
{
  ea.NX = ha.NX = c->NX;
  ea.NY = ha.NY = c->NY;
  ea.NZ = ha.NZ = c->NZ;
  ea.PX = ha.PX = c->PX;
  ea.plane = ha.plane = c->plane;
  for (int d = 0; d < 3; ++d) {
    ea.E[d] = P<T>(c->F[d]);
    ea.H[d] = P<T>(c->F[3 + d]);
    ha.H[d] = P<T>(c->F[3 + d]);
    ha.E[d] = P<T>(c->F[d]);
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
}

// Shell of a fast box inside [0, N-1]^3: up to six boxes (two x slabs of
// full y/z extent, two y slabs over the box's x range, two z slabs over its
// x-y range); the whole allocation when the box is empty.
static ShellBoxes shell_of(const Ctx *c, const int *fb)
{
  ShellBoxes sb{};
  const int N[3] = {c->NX, c->NY, c->NZ};
  auto add = [&](int x0, int x1, int y0, int y1, int z0, int z1) {
    if (x1 < x0 || y1 < y0 || z1 < z0) return;
    int *b = sb.box[sb.nb];
    b[0] = x0; b[1] = x1; b[2] = y0; b[3] = y1; b[4] = z0; b[5] = z1;
    sb.start[sb.nb + 1] = sb.start[sb.nb] + (long long)(x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
    ++sb.nb;
  };
  if (fb[0] > fb[1] || fb[2] > fb[3] || fb[4] > fb[5]) {
    add(0, N[0] - 1, 0, N[1] - 1, 0, N[2] - 1);
    return sb;
  }
  add(0, fb[0] - 1, 0, N[1] - 1, 0, N[2] - 1);
  add(fb[1] + 1, N[0] - 1, 0, N[1] - 1, 0, N[2] - 1);
  add(fb[0], fb[1], 0, fb[2] - 1, 0, N[2] - 1);
  add(fb[0], fb[1], fb[3] + 1, N[1] - 1, 0, N[2] - 1);
  add(fb[0], fb[1], fb[2], fb[3], 0, fb[4] - 1);
  add(fb[0], fb[1], fb[2], fb[3], fb[5] + 1, N[2] - 1);
  return sb;
}

// Once per context (at finalize): the sweeps' argument blocks go to device
// memory, where the shell kernels and the fused Debye step read them.
template <typename T> int prepare_sweeps(Ctx *c)
{
  // The fast box and its shell must tile the allocation exactly once, with
  // every box inside it (the shell kernels index from these boxes).
  for (int e = 0; e < 2; ++e) {
    const int *fb = e == 0 ? c->fastE : c->fastH;
    const ShellBoxes sb = shell_of(c, fb);
    long long n = 0;
    if (fb[0] <= fb[1] && fb[2] <= fb[3] && fb[4] <= fb[5]) {
      if (fb[0] < 1 || fb[2] < 1 || fb[4] < 1 || fb[1] > c->NX - 2 || fb[3] > c->NY - 2 || fb[5] > c->NZ - 2)
        return set_error(c, VGPU_EINVAL, "fast box not inside the grid");
      n = (long long)(fb[1] - fb[0] + 1) * (fb[3] - fb[2] + 1) * (fb[5] - fb[4] + 1);
    }
    const int N[3] = {c->NX, c->NY, c->NZ};
    for (int b = 0; b < sb.nb; ++b)
      for (int a = 0; a < 3; ++a)
        if (sb.box[b][2 * a] < 0 || sb.box[b][2 * a + 1] >= N[a] || sb.box[b][2 * a] > sb.box[b][2 * a + 1])
          return set_error(c, VGPU_EINVAL, "shell box outside the grid");
    if (n + sb.start[sb.nb] != (long long)c->NX * c->NY * c->NZ)
      return set_error(c, VGPU_EINVAL, "fast box and shell do not tile the grid");
  }
  EArgs<T> ea{};
  HArgs<T> ha{};
  fill_volume_args<T>(c, ea, ha);
  c->d_eargs = dev_alloc(c, sizeof(ea));
  c->d_hargs = dev_alloc(c, sizeof(ha));
  if (!c->d_eargs || !c->d_hargs) return set_error(c, VGPU_ENOMEM, "sweep arguments");
  if (hipMemcpy(c->d_eargs, &ea, sizeof(ea), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_hargs, &ha, sizeof(ha), hipMemcpyHostToDevice) != hipSuccess)
    return set_error(c, VGPU_EDEVICE, "sweep arguments upload");
  return VGPU_OK;
}

template <typename T> int run_steps(Ctx *c, uint64_t nsteps, hipStream_t s, bool timeKernels)
{
  const double pi = 4.0 * atan(1.0);
  const double dt = c->grid.dt;
  timeKernels = timeKernels && !c->evE.empty();
  EArgs<T> ea{};
  HArgs<T> ha{};
  fill_volume_args<T>(c, ea, ha);
  // Fast kernels: 64x4 tiles over the fast boxes, marching z in chunks sized
  // for about 16k blocks at 272^3 (at least 4 planes so the carry pays).
  const dim3 vblk(64, 4, 1);
  auto fast_grid = [&](const int *fb, int &kchunk) {
    if (fb[0] > fb[1] || fb[2] > fb[3] || fb[4] > fb[5]) return dim3(0, 0, 0);
    const int gx = (fb[1] - fb[0] + 64) / 64, gy = (fb[3] - fb[2] + 4) / 4, nz = fb[5] - fb[4] + 1;
    kchunk = (int)((long long)nz * gx * gy / 16384);
    kchunk = std::max(4, std::min(kchunk, 32));
    return dim3(gx, gy, (nz + kchunk - 1) / kchunk);
  };
  FastE<T> fe{};
  FastH<T> fh{};
  const dim3 egrid = fast_grid(c->fastE, fe.kchunk), hgrid = fast_grid(c->fastH, fh.kchunk);
  fe.NX = fh.NX = c->NX; fe.NY = fh.NY = c->NY; fe.NZ = fh.NZ = c->NZ; fe.PX = fh.PX = c->PX;
  fe.plane = fh.plane = c->plane;
  for (int q = 0; q < 6; ++q) { fe.box[q] = c->fastE[q]; fh.box[q] = c->fastH[q]; }
  fe.nchx = egrid.x; fh.nchx = hgrid.x;
  for (int d = 0; d < 3; ++d) {
    fe.E[d] = ea.E[d]; fe.H[d] = ea.H[d]; fe.idh[d] = ea.idh[d]; fe.alpha[d] = ea.alpha[d]; fe.beta[d] = ea.beta[d];
    fh.H[d] = ha.H[d]; fh.E[d] = ha.E[d]; fh.ide[d] = ha.ide[d]; fh.gamma[d] = ha.gamma[d];
  }
  fe.ecls = ea.ecls; fe.etab = ea.etab; fe.dfuse = ea.dfuse; fe.ap = P<EArgs<T>>(c->d_eargs);
  fh.hcls = ha.hcls; fh.htab = ha.htab;
  const ShellBoxes esh = shell_of(c, c->fastE), hsh = shell_of(c, c->fastH);
  const EArgs<T> *eap = P<EArgs<T>>(c->d_eargs);
  const HArgs<T> *hap = P<HArgs<T>>(c->d_hargs);
  auto fast_e = [&]() {
    if (egrid.x) {
      if (c->indexed) hipLaunchKernelGGL((k_e_fast<T, true>), egrid, vblk, 0, s, fe);
      else hipLaunchKernelGGL((k_e_fast<T, false>), egrid, vblk, 0, s, fe);
    }
  };
  auto shell_e = [&]() {
    if (esh.nb) {
      const dim3 g((unsigned)((esh.start[esh.nb] + 255) / 256));
      if (c->indexed) hipLaunchKernelGGL((k_e_shell<T, true>), g, dim3(256), 0, s, esh, eap);
      else hipLaunchKernelGGL((k_e_shell<T, false>), g, dim3(256), 0, s, esh, eap);
    }
  };
  auto fast_h = [&]() {
    if (hgrid.x) {
      if (c->indexed) hipLaunchKernelGGL((k_h_fast<T, true>), hgrid, vblk, 0, s, fh);
      else hipLaunchKernelGGL((k_h_fast<T, false>), hgrid, vblk, 0, s, fh);
    }
  };
  auto shell_h = [&]() {
    if (hsh.nb) {
      const dim3 g((unsigned)((hsh.start[hsh.nb] + 255) / 256));
      if (c->indexed) hipLaunchKernelGGL((k_h_shell<T, true>), g, dim3(256), 0, s, hsh, hap);
      else hipLaunchKernelGGL((k_h_shell<T, false>), g, dim3(256), 0, s, hsh, hap);
    }
  };

  MurArgs<T> ma{};
  int murmax1 = 1, murmax2 = 1;
  if (c->murmask) {
    ma.NX = c->NX; ma.NY = c->NY; ma.PX = c->PX; ma.plane = c->plane;
    for (int d = 0; d < 3; ++d) {
      ma.E[d] = P<T>(c->F[d]); ma.H[d] = P<T>(c->F[3 + d]);
      ma.alpha[d] = P<T>(c->alpha[d]); ma.beta[d] = P<T>(c->beta[d]); ma.idh[d] = P<T>(c->idh[d]);
    }
    static const int order[6][2] = {{1, 2}, {1, 2}, {0, 2}, {0, 2}, {1, 0}, {1, 0}};
    for (int f = 0; f < 6; ++f) {
      ma.zeta[f] = T(c->zeta[f]);
      for (int q = 0; q < 2; ++q) {
        const int slot = 2 * f + q;
        if ((c->murmask >> f) & 1u) {
          const int comp = order[f][q];
          ma.comp[slot] = comp;
          for (int e = 0; e < 6; ++e) ma.box[slot][e] = c->fmlim[f][comp][e];
          const int ax = f >> 1, a1 = ax == 0 ? 1 : 0, a2 = ax == 2 ? 1 : 2;
          murmax1 = std::max(murmax1, ma.box[slot][2 * a1 + 1] - ma.box[slot][2 * a1] + 1);
          murmax2 = std::max(murmax2, ma.box[slot][2 * a2 + 1] - ma.box[slot][2 * a2] + 1);
        } else {
          ma.comp[slot] = -1;
        }
      }
    }
  }

  DebyeArgs<T> da{};
  if (c->debyeSites) {
    da.comp = P<int>(c->d_dcomp); da.cell = P<long long>(c->d_dcell); da.medium = P<int>(c->d_dmed);
    for (int d = 0; d < 3; ++d) { da.E[d] = P<T>(c->F[d]); da.beta[d] = P<T>(c->beta[d]); }
    da.lastE = P<T>(c->d_dlast); da.J = P<T>(c->d_dJ); da.stride = c->debyeSites;
    da.dalpha = P<double>(c->d_dalpha); da.dbeta = P<double>(c->d_dbeta); da.npoles = P<int>(c->d_dnp);
    da.maxPoles = c->debyeMaxPoles; da.cplx = c->debyeCplx;
    if (c->indexed) { da.ecls = P<uint16_t>(c->d_ecls); da.etab = P<T>(c->d_etab); }
  }

  SibcEArgs<T> se{};
  SibcHArgs<T> sh{};
  if (c->sibcFaces) {
    se.nfaces = c->sibcFaces; se.ijk = P<int>(c->d_sijk); se.surf = P<int>(c->d_ssurf);
    se.adjA = P<unsigned char>(c->d_sadjA); se.adjB = P<unsigned char>(c->d_sadjB);
    se.normal = P<int>(c->d_snormal); se.A = P<double>(c->d_sA); se.B = P<double>(c->d_sB);
    se.filtset = P<int>(c->d_sset); se.npoles = P<int>(c->d_snp); se.asymp = P<double>(c->d_sasymp);
    se.q = P<double>(c->d_sq); se.maxPoles = c->sibcMaxPoles; se.zeta = P<double>(c->d_szeta);
    se.old = P<double>(c->d_sold); se.Etan = P<T>(c->d_setan);
    for (int d = 0; d < 3; ++d) { se.E[d] = P<T>(c->F[d]); se.H[d] = P<T>(c->F[3 + d]); }
    se.NY = c->NY; se.PX = c->PX; se.plane = c->plane;
    sh.nrows = c->sibcHRows; sh.rowcomp = P<int>(c->d_shcomp); sh.rowcell = P<long long>(c->d_shcell);
    sh.rowstart = P<int>(c->d_shstart); sh.rowzero = P<int>(c->d_shzero);
    sh.ent = P<SibcHEntry>(c->d_shent); sh.Etan = P<T>(c->d_setan);
    for (int d = 0; d < 3; ++d) { sh.H[d] = P<T>(c->F[3 + d]); sh.gamma[d] = P<T>(c->gamma[d]); }
  }

  CorrArgs<T> ce[2];
  for (int e = 0; e < 2; ++e) {
    CorrArgs<T> &a = ce[e];
    a.nrows = c->corrRows[e];
    a.rowcomp = P<int>(c->d_ccomp[e]); a.rowcell = P<long long>(c->d_ccell[e]);
    a.rowstart = P<int>(c->d_cstart[e]); a.ent = P<CorrEntry>(c->d_cent[e]);
    for (int d = 0; d < 3; ++d) {
      a.F[d] = P<T>(c->F[3 * (1 - e) + d]);
      a.alpha[d] = P<T>(c->alpha[d]);
      a.coef[d] = e == 0 ? P<T>(c->beta[d]) : P<T>(c->gamma[d]);
    }
    a.electric = e == 0;
    a.wf = P<WaveformDev>(c->d_wf); a.src = P<SourceDev>(c->d_src); a.pw = P<PwDev>(c->d_pw);
    a.aux = P<T>(c->d_aux); a.auxmax = c->auxmax; a.eta0 = c->grid.eta0;
  }
  // the E list is built from E components, the H list from H components
  for (int d = 0; d < 3; ++d) { ce[0].F[d] = P<T>(c->F[d]); ce[1].F[d] = P<T>(c->F[3 + d]); }

  AuxArgs<T> xa{};
  bool anyAux = false;
  for (auto &p : c->pw) anyAux |= p.d.use_aux != 0;
  if (anyAux) {
    xa.npw = (int)c->pw.size(); xa.pw = P<PwDev>(c->d_pw); xa.aux = P<T>(c->d_aux);
    xa.pml = P<T>(c->d_auxpml); xa.prof = P<T>(c->d_auxprof); xa.auxmax = c->auxmax;
    xa.wf = P<WaveformDev>(c->d_wf);
  }

  ObsArgs<T> oa{};
  int maxnf = 1;
  if (!c->pobs.empty()) {
    oa.nobs = (int)c->pobs.size(); oa.cell = P<long long>(c->d_ocell);
    oa.first = P<unsigned long long>(c->d_ofirst); oa.last = P<unsigned long long>(c->d_olast);
    oa.nfreq = P<int>(c->d_onf); oa.dftoff = P<long long>(c->d_ooff); oa.omega = P<double>(c->d_oomega);
    oa.td = P<double>(c->d_otd); oa.tdsteps = c->totalSteps; oa.dre = P<double>(c->d_odre);
    oa.dim = P<double>(c->d_odim);
    for (int d = 0; d < 6; ++d) oa.F[d] = P<T>(c->F[d]);
    for (auto &o : c->pobs) maxnf = std::max(maxnf, (int)o.omega.size());
  }
  VolObsArgs<T> va{};
  if (c->vnodes) {
    va.nodecell = P<long long>(c->d_vcell); va.nodeobs = P<int>(c->d_vobs); va.nnodes = c->vnodes;
    va.first = P<unsigned long long>(c->d_vfirst); va.last = P<unsigned long long>(c->d_vlast);
    va.stage = P<float>(c->d_vstage);
    for (int d = 0; d < 6; ++d) va.F[d] = P<T>(c->F[d]);
  }

  auto ghost = [&](int electric) {
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
      for (int q = 0; q < nc; ++q) {
        g.F = P<T>(c->F[comps[q]]);
        hipLaunchKernelGGL(k_ghost<T>, gg, gb, 0, s, g);
      }
    }
  };

  // Time-dependent kernels read the step number from the device counter
  // (step n = counter + offset within this run), so the same launches can be
  // captured once in a graph and replayed for later runs.
  unsigned long long *stepp = P<unsigned long long>(c->d_step);
  oa.step = stepp; oa.dt = dt;
  va.step = stepp;
  xa.step = stepp; xa.dt = dt;
  for (int e = 0; e < 2; ++e) { ce[e].step = stepp; ce[e].dt = dt; ce[e].half = e == 0 ? 0.0 : 0.5; }
  int ne = 0, nh = 0;
  for (uint64_t q = 0; q < nsteps; ++q) {
    const int soff = (int)q;
    if (oa.nobs) {
      oa.soff = soff;
      hipLaunchKernelGGL(k_observe<T>, dim3((maxnf + 63) / 64, oa.nobs), dim3(64), 0, s, oa);
    }
    if (va.nnodes) {
      va.soff = soff;
      hipLaunchKernelGGL(k_observe_volume<T>, dim3((unsigned)((va.nnodes + 255) / 256)), dim3(256), 0, s, va);
    }
    if (c->murmask)
      hipLaunchKernelGGL(k_mur_e<T>, dim3((murmax1 + 15) / 16, (murmax2 + 15) / 16, 12), dim3(16, 16), 0, s, ma);
    // events bracket the fast kernels (the dominant launches)
    if (timeKernels && ne < (int)c->evE.size() / 2) hipEventRecord(c->evE[2 * ne], s);
    fast_e();
    if (timeKernels && ne < (int)c->evE.size() / 2) { hipEventRecord(c->evE[2 * ne + 1], s); ++ne; }
    shell_e();
    for (int l = c->debyeFused ? 1 : 0; l < c->debyeLayers; ++l) {
      da.base = c->debyeLayerStart[l];
      da.n = c->debyeLayerStart[l + 1] - c->debyeLayerStart[l];
      if (da.n == 0) continue;
      hipLaunchKernelGGL(k_debye<T>, dim3((unsigned)((da.n + 255) / 256)), dim3(256), 0, s, da);
    }
    if (c->sibcFaces)
      hipLaunchKernelGGL(k_sibc_e<T>, dim3((unsigned)((se.nfaces + 127) / 128)), dim3(128), 0, s, se);
    if (anyAux) {
      xa.soff = soff;
      hipLaunchKernelGGL(k_aux_e<T>, dim3(xa.npw), dim3(256), 0, s, xa, pi);
    }
    if (ce[0].nrows) {
      ce[0].soff = soff;
      hipLaunchKernelGGL(k_corr<T>, dim3((unsigned)((ce[0].nrows + 255) / 256)), dim3(256), 0, s, ce[0], pi);
    }
    ghost(1);
    if (timeKernels && nh < (int)c->evH.size() / 2) hipEventRecord(c->evH[2 * nh], s);
    fast_h();
    if (timeKernels && nh < (int)c->evH.size() / 2) { hipEventRecord(c->evH[2 * nh + 1], s); ++nh; }
    shell_h();
    if (c->sibcFaces)
      hipLaunchKernelGGL(k_sibc_h<T>, dim3((unsigned)((sh.nrows + 255) / 256)), dim3(256), 0, s, sh);
    if (anyAux) hipLaunchKernelGGL(k_aux_h<T>, dim3(xa.npw), dim3(256), 0, s, xa);
    if (ce[1].nrows) {
      ce[1].soff = soff;
      hipLaunchKernelGGL(k_corr<T>, dim3((unsigned)((ce[1].nrows + 255) / 256)), dim3(256), 0, s, ce[1], pi);
    }
    ghost(0);
  }
  hipLaunchKernelGGL(k_advance_step, dim3(1), dim3(64), 0, s, stepp, (unsigned long long)nsteps);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return set_error(c, VGPU_EDEVICE, "kernel launch failed: %s", hipGetErrorString(err));
  return VGPU_OK;
}

}  // namespace vgpu
