// Per-precision time-step driver: launches one step's kernels in the
// reference's order (simulation.c:157-222):
//   observers -> Mur E -> E sweep (inner grid + UPML) -> Debye -> SIBC E ->
//   aux grids E -> sources/TF-SF E -> ghost E -> H sweep (inner grid + Mur
//   normal H + UPML) -> SIBC H -> aux grids H -> sources/TF-SF H -> ghost H
#pragma once
#include "vgpu_kernels.h"

namespace vgpu {

template <typename T> static T *P(void *p) { return static_cast<T *>(p); }

template <typename T> static void fill_volume_args(Ctx *c, EArgs<T> &ea, HArgs<T> &ha, int kchunk)
{
  ea.NX = ha.NX = c->NX;
  ea.NY = ha.NY = c->NY;
  ea.NZ = ha.NZ = c->NZ;
  ea.PX = ha.PX = c->PX;
  ea.plane = ha.plane = c->plane;
  ea.k0 = ha.k0 = 0;
  ea.kchunk = ha.kchunk = kchunk;
  ea.x0 = ha.x0 = c->X0;
  for (int q = 0; q < 6; ++q) { ea.fast[q] = c->fastE[q]; ha.fast[q] = c->fastH[q]; }
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
    ea.dbase = P<int>(c->d_dbase); ea.NCH = c->NCH;
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

template <typename T> int run_steps(Ctx *c, uint64_t nsteps, hipStream_t s, bool timeKernels)
{
  const double pi = 4.0 * atan(1.0);
  const double dt = c->grid.dt;
  timeKernels = timeKernels && !c->evE.empty();
  // Volume launch: 64x4 tiles, x chunks aligned to the fast box.
  const dim3 vblk(64, 4, 1);
  const int gx = c->NCH, gy = (c->NY + 3) / 4;
  // z planes per block: short marches (about 16k blocks at 272^3) keep many
  // rows in flight; at least 4 planes so the register carry pays.
  int kchunk = (int)((long long)c->NZ * gx * gy / 16384);
  if (kchunk < 4) kchunk = 4;
  if (kchunk > 32) kchunk = 32;
  const dim3 vgrid(gx, gy, (c->NZ + kchunk - 1) / kchunk);
  EArgs<T> ea;
  HArgs<T> ha;
  fill_volume_args<T>(c, ea, ha, kchunk);

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
    if (timeKernels && ne < (int)c->evE.size() / 2) hipEventRecord(c->evE[2 * ne], s);
    if (c->indexed) hipLaunchKernelGGL((k_volume_e<T, true>), vgrid, vblk, 0, s, ea);
    else hipLaunchKernelGGL((k_volume_e<T, false>), vgrid, vblk, 0, s, ea);
    if (timeKernels && ne < (int)c->evE.size() / 2) { hipEventRecord(c->evE[2 * ne + 1], s); ++ne; }
    for (int l = c->debyeFused ? 1 : 0; l < c->debyeLayers; ++l) {
      da.base = c->debyeLayerStart[l];
      da.n = c->debyeLayerStart[l + 1] - c->debyeLayerStart[l];
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
    if (c->indexed) hipLaunchKernelGGL((k_volume_h<T, true>), vgrid, vblk, 0, s, ha);
    else hipLaunchKernelGGL((k_volume_h<T, false>), vgrid, vblk, 0, s, ha);
    if (timeKernels && nh < (int)c->evH.size() / 2) { hipEventRecord(c->evH[2 * nh + 1], s); ++nh; }
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
