/*
 * Waveform evaluation and observer output in the reference's file formats
 * (waveform.c:413-805, observer.c:622-1198).
 *
 * Point observers write eh_<name>_td.asc / eh_<name>_fd.asc, every waveform
 * gets wf_<name>_td.asc / wf_<name>_fd.asc, volume observers write the
 * binary impulse.dat with excite.dat and process.dat beside it.  Values are
 * formatted with %16.8e like the reference, or %24.16e when full precision is
 * requested (the FP64 comparison needs all digits).
 */
#include "vh_internal.h"

/* ---- waveforms ---- */

static real gaussian_pulse(real time, real width) { return exp(-0.5 * (pow(time / width, 2.0))); }

static real diff_gaussian_pulse(real time, real width)
{
  return -time / width * exp(-0.5 * (pow(time / width, 2.0)));
}

static real ricker_wavelet(real time, real width)
{
  return (1.0 - pow(time / width, 2)) * exp(-0.5 * (pow(time / width, 2.0)));
}

static real mod_gaussian_pulse(const vh_solve *s, real time, real width, real frequency)
{
  return exp(-0.5 * (pow(time / width, 2.0))) * sin(2.0 * s->pi * frequency * time);
}

static real compact_envelope(const vh_solve *s, real time, real width, real span)
{
  real pi = s->pi;
  if (time <= 0.0) return 0.0;
  if (time < span * width)
    return 1.0 / 32.0 * (10.0 - 15.0 * cos(pi / width * time) + 6.0 * cos(2.0 * pi / width * time) -
                         cos(3.0 * pi / width * time));
  return span < 1.5 ? 1.0 : 0.0;
}

static real diff_compact_pulse(const vh_solve *s, real time, real width)
{
  real pi = s->pi;
  if (time <= 0.0) return 0.0;
  if (time < 2.0 * width)
    return 1.0 / 32.0 * (15.0 * sin(pi / width * time) - 12.0 * sin(2.0 * pi / width * time) +
                         3.0 * sin(3.0 * pi / width * time));
  return 0.0;
}

/* Cubic spline interpolation with the last-used interval cached
 * (waveform.c:762-805). */
static double eval_spline(vh_waveform *w, real time)
{
  real *T = w->table;
  unsigned long n = w->tableSize, k, klo, khi;
  real h, b, a;
  klo = w->lastIdx;
  khi = klo + 1UL;
  if (klo >= n - 1UL) klo = 0;
  if (T[3 * klo] > time || T[3 * (klo + 1)] < time) {
    klo = 0;
    khi = n - 1UL;
    while (khi - klo > 1UL) {
      k = (khi + klo) >> 1;
      if (T[3 * k] > time) khi = k;
      else klo = k;
    }
  }
  w->lastIdx = klo;
  h = T[3 * khi] - T[3 * klo];
  a = (T[3 * khi] - time) / h;
  b = (time - T[3 * klo]) / h;
  return a * T[3 * klo + 1] + b * T[3 * khi + 1] +
         ((a * a * a - a) * T[3 * klo + 2] + (b * b * b - b) * T[3 * khi + 2]) * (h * h) / 6.0;
}

real vh_waveform_value(vh_solve *s, real t, int waveform, real delay)
{
  vh_waveform *w = &s->waveforms[waveform];
  real tt = t - delay - w->delay;
  switch (w->type) {
  case WT_GAUSSIAN_PULSE:
  case WT_NARROW_GAUSSIAN_PULSE: return w->size * gaussian_pulse(tt, w->width);
  case WT_DIFFERENTIATED_GAUSSIAN_PULSE: return w->size * diff_gaussian_pulse(tt, w->width);
  case WT_RICKER_WAVELET: return w->size * ricker_wavelet(tt, w->width);
  case WT_MODULATED_GAUSSIAN_PULSE: return w->size * mod_gaussian_pulse(s, tt, w->width, w->frequency);
  case WT_COMPACT_PULSE: return w->size * compact_envelope(s, tt, w->width, 2.0);
  case WT_DIFFERENTIATED_COMPACT_PULSE: return w->size * diff_compact_pulse(s, tt, w->width);
  case WT_MODULATED_COMPACT_PULSE: {
    real env = compact_envelope(s, tt, w->width, 2.0);
    return w->size * (real)(env * sin(2.0 * s->pi * w->frequency * tt));
  }
  case WT_RAMPED_SINUSOID: {
    real ramp = compact_envelope(s, tt, w->width, 1.0);
    return w->size * (real)(ramp * sin(2.0 * s->pi * w->frequency * tt));
  }
  case WT_EXTERNAL: {
    real v;
    if (tt < w->table[0] || tt > w->table[3 * (w->tableSize - 1)]) v = 0.0;
    else v = eval_spline(w, tt);
    return w->size * v;
  }
  default: return 0.0;
  }
}

/* ---- observers ---- */

int vh_num_point_observers(const vh_solve *s, int domain)
{
  int n = 0;
  for (int o = 0; o < s->numObs; o++)
    if (s->obs[o].format == OF_ASCII && s->obs[o].quantity == OQ_EH && s->obs[o].domain == domain) n++;
  return n;
}

static const char *vfmt(const vh_solve *s) { return s->fullPrecision ? "%24.16e " : "%16.8e "; }

static void write_process_dat(vh_solve *s)
{
  FILE *fp = vh_open_output(s, "process.dat", "w");
  if (!fp) vh_fail(s, "failed to open process.dat");
  real fstart, fstop, fstep, meshSize;
  meshSize = s->dt * (2.0 * s->c0);
  fstart = 0.0;
  fstep = 1.0 / (s->numTimeSteps * s->dt);
  fstop = s->numTimeSteps * fstep;
  char comment[70];
  snprintf(comment, sizeof(comment) - 1, "%s", s->comment);
  fprintf(fp, "CE %s\n", comment);
  unsigned long nbin = 0;
  for (int o = 0; o < s->numObs; o++) nbin += s->obs[o].format == OF_BINARY;
  fprintf(fp, "%lu\n", nbin);
  int first = -1;
  for (int o = 0; o < s->numObs; o++) {
    vh_observer *ob = &s->obs[o];
    if (ob->domain != OD_TIME || ob->format != OF_BINARY) continue;
    if (first < 0) first = o;
    fprintf(fp, " %d %d %d %d %d %d %d %d %d\n", ob->gbbox[XLO] - s->gibox[XLO], ob->gbbox[XHI] - s->gibox[XLO],
            ob->step[0], ob->gbbox[YLO] - s->gibox[YLO], ob->gbbox[YHI] - s->gibox[YLO], ob->step[1],
            ob->gbbox[ZLO] - s->gibox[ZLO], ob->gbbox[ZHI] - s->gibox[ZLO], ob->step[2]);
  }
  fprintf(fp, "%lu %lu\n", s->startTimeStep + 1, s->stopTimeStep + 1);
  fprintf(fp, "%g %g %g\n", fstart, fstop, fstep);
  fprintf(fp, "%g\n", meshSize);
  if (first >= 0) {
    vh_observer *ob = &s->obs[first];
    fprintf(fp, " %d %d %d %d %d %d %d\n", ob->gbbox[XLO] - s->gibox[XLO], ob->gbbox[XHI] - s->gibox[XLO],
            ob->gbbox[YLO] - s->gibox[YLO], ob->gbbox[YHI] - s->gibox[YLO], ob->gbbox[ZLO] - s->gibox[ZLO],
            ob->gbbox[ZHI] - s->gibox[ZLO], EX + 1);
  }
  fclose(fp);
}

int vh_open_observers(vh_solve *s)
{
  jmp_buf jb;
  if (setjmp(jb)) {
    s->jb = NULL;
    return -1;
  }
  s->jb = &jb;
  char fileName[VH_PATH];
  for (int o = 0; o < s->numObs; o++) {
    vh_observer *ob = &s->obs[o];
    if (ob->format != OF_ASCII) continue;
    const char *pre = ob->quantity == OQ_WF ? "wf" : "eh";
    const char *suf = ob->domain == OD_TIME ? "td" : "fd";
    snprintf(fileName, sizeof(fileName), "%s_%s_%s.asc", pre, ob->name, suf);
    ob->outputFile = vh_open_output(s, fileName, "w");
    if (!ob->outputFile) vh_fail(s, "failed to open output file %s", fileName);
    if (ob->quantity == OQ_WF) {
      fprintf(ob->outputFile, "# Waveform# %lu\n", (unsigned long)ob->waveform);
      if (ob->domain == OD_TIME)
        fprintf(ob->outputFile, "# %6s %16s %16s\n", "ts (-)", "t (s)", "wf (-)");
      else
        fprintf(ob->outputFile, "# %14s %16s %16s\n", "f (Hz)", "Re(wf) (-)", "Im(wf) (-)");
    } else {
      real phys[3] = {s->xlines[ob->mbbox[XLO]], s->ylines[ob->mbbox[YLO]], s->zlines[ob->mbbox[ZLO]]};
      fprintf(ob->outputFile, "# (%d,%d,%d)->(%g,%g,%g)\n", ob->mbbox[XLO], ob->mbbox[YLO], ob->mbbox[ZLO],
              phys[0], phys[1], phys[2]);
      if (ob->domain == OD_TIME)
        fprintf(ob->outputFile, "# %6s %16s %16s %16s %16s %16s %16s %16s\n", "ts (-)", "t (s)", "Ex (V/m)",
                "Ey (V/m)", "Ez (V/m)", "Hx (A/m)", "Hy (A/m)", "Hz (A/m)");
      else
        fprintf(ob->outputFile, "# %14s %16s %16s %16s %16s %16s %16s %16s %16s %16s %16s %16s %16s\n", "f (Hz)",
                "Re(Ex) (V/m)", "Im(Ex) (V/m)", "Re(Ey) (V/m)", "Im(Ey) (V/m)", "Re(Ez) (V/m)", "Im(Ez) (V/m)",
                "Re(Hx) (V/m)", "Im(Hx) (V/m)", "Re(Hy) (V/m)", "Im(Hy) (V/m)", "Re(Hz) (V/m)", "Im(Hz) (V/m)");
    }
  }
  int nbin = 0;
  for (int o = 0; o < s->numObs; o++) nbin += s->obs[o].format == OF_BINARY;
  if (nbin) {
    s->impulseFile = vh_open_output(s, "impulse.dat", "wb");
    if (!s->impulseFile) vh_fail(s, "failed to open impulse.dat");
    int n = (int)s->numOutTimeSteps;
    float dtf = (float)s->obsDt;
    fwrite(&n, sizeof(int), 1, s->impulseFile);
    fwrite(&dtf, sizeof(float), 1, s->impulseFile);
    s->exciteFile = vh_open_output(s, "excite.dat", "w");
    if (!s->exciteFile) vh_fail(s, "failed to open excite.dat");
    fprintf(s->exciteFile, " %d\n", 8);
    write_process_dat(s);
  }
  s->jb = NULL;
  return 0;
}

/* Read the six components at a node from the host arrays. */
void vh_observer_point_values(const vh_solve *s, double *vals)
{
  for (int o = 0; o < s->numObs; o++) {
    const vh_observer *ob = &s->obs[o];
    if (ob->format != OF_ASCII || ob->quantity != OQ_EH) continue;
    size_t id = VH_IDX(s, ob->gbbox[XLO], ob->gbbox[YLO], ob->gbbox[ZLO]);
    const real *F[6] = {s->Ex, s->Ey, s->Ez, s->Hx, s->Hy, s->Hz};
    for (int c = 0; c < 6; c++) vals[6 * o + c] = F[c][id];
  }
}

static void write_td_line(vh_solve *s, FILE *fp, unsigned long n, real t, const real *v, int nc)
{
  const char *f = vfmt(s);
  if (s->fullPrecision) fprintf(fp, "%8lu %24.16e ", n, t);
  else fprintf(fp, "%8lu %16.8e ", n, t);
  for (int c = 0; c < nc; c++) fprintf(fp, f, v[c]);
  fprintf(fp, "\n");
}

static void accumulate_dft(vh_solve *s, vh_observer *ob, real t, const real *v)
{
  for (int c = 0; c < ob->numComp; c++)
    for (unsigned long f = 0; f < s->numFreq; f++) {
      real omegat = s->omega[f] * t;
      ob->dft_re[c * s->numFreq + f] += v[c] * cos(omegat);
      ob->dft_im[c * s->numFreq + f] -= v[c] * sin(omegat);
    }
}

/* Step n's excite.dat line and, inside the output window, its impulse.dat
 * rows (observer.c:1164-1198): every binary observer's nodes in k, j, i
 * order, each as three int indices and six float field values, from binVals
 * (six floats per node in file order) or, when NULL, the host arrays. */
static void write_binary_step(vh_solve *s, unsigned long n, real t, const float *binVals)
{
  int nbin = 0;
  for (int o = 0; o < s->numObs; o++) nbin += s->obs[o].format == OF_BINARY;
  if (!nbin) return;
  if (s->exciteFile) fprintf(s->exciteFile, "%16.8e\n", vh_waveform_value(s, t, 0, 0.0));
  if (n < s->startTimeStep || n > s->stopTimeStep) return;
  const real *F[6] = {s->Ex, s->Ey, s->Ez, s->Hx, s->Hy, s->Hz};
  size_t bpos = 0;
  for (int o = 0; o < s->numObs; o++) {
    vh_observer *ob = &s->obs[o];
    if (ob->quantity == OQ_WF || ob->format != OF_BINARY) continue;
    for (int k = ob->gbbox[ZLO]; k <= ob->gbbox[ZHI]; k += ob->step[2])
      for (int j = ob->gbbox[YLO]; j <= ob->gbbox[YHI]; j += ob->step[1])
        for (int i = ob->gbbox[XLO]; i <= ob->gbbox[XHI]; i += ob->step[0]) {
          float out[6];
          if (binVals) {
            for (int c = 0; c < 6; c++) out[c] = binVals[6 * bpos + c];
          } else {
            size_t id = VH_IDX(s, i, j, k);
            for (int c = 0; c < 6; c++) out[c] = F[c][id];
          }
          bpos++;
          int io[3] = {i - s->gibox[XLO], j - s->gibox[YLO], k - s->gibox[ZLO]};
          if (s->impulseFile) {
            fwrite(io, sizeof(int), 3, s->impulseFile);
            fwrite(out, sizeof(float), 6, s->impulseFile);
          }
        }
  }
}

/* The binary part of a step on its own, for callers that stream volume
 * samples as runs complete (vh_observe_step then skips it). */
void vh_observe_binary_step(vh_solve *s, unsigned long n, real t, const float *binVals)
{
  write_binary_step(s, n, t, binVals);
}

/*
 * One call per time step, before the step's updates (simulation.c:167).
 * pointVals: [numObs][6] node values for EH point observers, or NULL to read
 * the host field arrays.  binVals: the binary observers' nodes, six floats
 * each, in file order, or NULL to read the host arrays.
 * The binary part is left out when the caller streams it
 * (s->binaryStreamed, vh_observe_binary_step).
 */
void vh_observe_step(vh_solve *s, unsigned long n, real t, const double *pointVals, const float *binVals)
{
  bool otValid = n >= s->startTimeStep && n <= s->stopTimeStep;
  if (!s->binaryStreamed) write_binary_step(s, n, t, binVals);
  const real *F[6] = {s->Ex, s->Ey, s->Ez, s->Hx, s->Hy, s->Hz};
  for (int o = 0; o < s->numObs; o++) {
    vh_observer *ob = &s->obs[o];
    real v[6];
    if (ob->quantity == OQ_WF) {
      v[0] = vh_waveform_value(s, t, ob->waveform, 0.0);
      if (ob->domain == OD_TIME) {
        if (s->openOutputs && ob->outputFile) write_td_line(s, ob->outputFile, n, t, v, 1);
      } else {
        accumulate_dft(s, ob, t, v);
      }
      continue;
    }
    if (!otValid || ob->format == OF_BINARY) continue;
    if (pointVals) {
      for (int c = 0; c < 6; c++) v[c] = pointVals[6 * o + c];
    } else {
      size_t id = VH_IDX(s, ob->gbbox[XLO], ob->gbbox[YLO], ob->gbbox[ZLO]);
      for (int c = 0; c < 6; c++) v[c] = F[c][id];
    }
    if (ob->domain == OD_TIME) {
      if (s->openOutputs && ob->outputFile) write_td_line(s, ob->outputFile, n, t, v, 6);
    } else if (!s->deviceDft) {
      accumulate_dft(s, ob, t, v);
    }
  }
}

/* Write DFTs normalised by the waveform DFT and close every file
 * (observer.c:653-699, 949-995). */
void vh_close_observers(vh_solve *s)
{
  const char *f = vfmt(s);
  for (int o = 0; o < s->numObs; o++) {
    vh_observer *ob = &s->obs[o];
    if (ob->format != OF_ASCII || ob->domain != OD_FREQ || !ob->outputFile) continue;
    for (unsigned long q = 0; q < s->numFreq; q++) {
      fprintf(ob->outputFile, f, s->startFreq + q * s->stepFreq);
      if (ob->quantity == OQ_WF) {
        fprintf(ob->outputFile, f, ob->dft_re[q]);
        fprintf(ob->outputFile, f, ob->dft_im[q]);
      } else {
        vh_observer *wo = &s->obs[ob->waveformObserver];
        real wf_r = wo->dft_re[q], wf_i = wo->dft_im[q];
        real denom = wf_r * wf_r + wf_i * wf_i;
        for (int c = 0; c < ob->numComp; c++) {
          real cr = ob->dft_re[c * s->numFreq + q], ci = ob->dft_im[c * s->numFreq + q];
          real re = (cr * wf_r + ci * wf_i) / denom;
          real im = (ci * wf_r - cr * wf_i) / denom;
          fprintf(ob->outputFile, f, re);
          fprintf(ob->outputFile, f, im);
        }
      }
      fprintf(ob->outputFile, "\n");
    }
  }
  for (int o = 0; o < s->numObs; o++)
    if (s->obs[o].outputFile) {
      fclose(s->obs[o].outputFile);
      s->obs[o].outputFile = NULL;
    }
  if (s->impulseFile) { fclose(s->impulseFile); s->impulseFile = NULL; }
  if (s->exciteFile) { fclose(s->exciteFile); s->exciteFile = NULL; }
}
