/*
 * cpu_ref - CPU restatement of the reference time-stepping loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is checked
 * against and the CPU baseline the benchmark times beside it.  Nothing in the
 * product library links or calls it.
 *
 * It runs propagate() (simulation.c:136-230) over a vh_solve built by the host
 * set-up library, in the reference's own k-fastest [i][j][k] layout, with the
 * same sweep order, the same operand order inside every update and the same
 * quirks (Mur E before the grid sweep, alpha applied again by the TF/SF
 * corrections, the Debye lastE cache, the SIBC Y-normal Exb term).  Volume
 * sweeps are OpenMP-parallel over i like the reference's.
 *
 * Compiled twice: -DVH_REAL=float reproduces the reference float build bit
 * for bit (pinned against it), -DVH_REAL=double is the FP64 oracle.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "vh.h"

#define IDX(i, j, k) VH_IDX(s, i, j, k)

/* Curl operators with per-line inverse lengths (grid.h:264-269). */
#define CURL_HX(s, id, j, k, SJ) (s->idhy[j] * (s->Hz[id] - s->Hz[(id) - (SJ)]) + s->idhz[k] * (s->Hy[(id) - 1] - s->Hy[id]))
#define CURL_HY(s, id, i, k, SI) (s->idhz[k] * (s->Hx[id] - s->Hx[(id) - 1]) + s->idhx[i] * (s->Hz[(id) - (SI)] - s->Hz[id]))
#define CURL_HZ(s, id, i, j, SI, SJ) (s->idhx[i] * (s->Hy[id] - s->Hy[(id) - (SI)]) + s->idhy[j] * (s->Hx[(id) - (SJ)] - s->Hx[id]))
#define CURL_EX(s, id, j, k, SJ) (s->idez[k] * (s->Ey[(id) + 1] - s->Ey[id]) + s->idey[j] * (s->Ez[id] - s->Ez[(id) + (SJ)]))
#define CURL_EY(s, id, i, k, SI) (s->idex[i] * (s->Ez[(id) + (SI)] - s->Ez[id]) + s->idez[k] * (s->Ex[id] - s->Ex[(id) + 1]))
#define CURL_EZ(s, id, i, j, SI, SJ) (s->idey[j] * (s->Ex[(id) + (SJ)] - s->Ex[id]) + s->idex[i] * (s->Ey[id] - s->Ey[(id) + (SI)]))

/* ---- inner grid (grid.c:709-888) ---- */

static void grid_e(vh_solve *s)
{
  const size_t SJ = (size_t)s->numCells[2], SI = SJ * (size_t)s->numCells[1];
  int (*L)[6] = s->gfilim;
#pragma omp parallel for
  for (int i = L[EX][XLO]; i <= L[EX][XHI]; i++)
    for (int j = L[EX][YLO]; j <= L[EX][YHI]; j++)
      for (int k = L[EX][ZLO]; k <= L[EX][ZHI]; k++) {
        size_t id = IDX(i, j, k);
        s->Ex[id] = s->alphaEx[id] * s->Ex[id] + s->betaEx[id] * CURL_HX(s, id, j, k, SJ);
      }
#pragma omp parallel for
  for (int i = L[EY][XLO]; i <= L[EY][XHI]; i++)
    for (int j = L[EY][YLO]; j <= L[EY][YHI]; j++)
      for (int k = L[EY][ZLO]; k <= L[EY][ZHI]; k++) {
        size_t id = IDX(i, j, k);
        s->Ey[id] = s->alphaEy[id] * s->Ey[id] + s->betaEy[id] * CURL_HY(s, id, i, k, SI);
      }
#pragma omp parallel for
  for (int i = L[EZ][XLO]; i <= L[EZ][XHI]; i++)
    for (int j = L[EZ][YLO]; j <= L[EZ][YHI]; j++)
      for (int k = L[EZ][ZLO]; k <= L[EZ][ZHI]; k++) {
        size_t id = IDX(i, j, k);
        s->Ez[id] = s->alphaEz[id] * s->Ez[id] + s->betaEz[id] * CURL_HZ(s, id, i, j, SI, SJ);
      }
}

static void h_box(vh_solve *s, const int L[6][6])
{
  const size_t SJ = (size_t)s->numCells[2], SI = SJ * (size_t)s->numCells[1];
#pragma omp parallel for
  for (int i = L[HX][XLO]; i <= L[HX][XHI]; i++)
    for (int j = L[HX][YLO]; j <= L[HX][YHI]; j++)
      for (int k = L[HX][ZLO]; k <= L[HX][ZHI]; k++) {
        size_t id = IDX(i, j, k);
        s->Hx[id] = s->Hx[id] + s->gammaHx[id] * CURL_EX(s, id, j, k, SJ);
      }
#pragma omp parallel for
  for (int i = L[HY][XLO]; i <= L[HY][XHI]; i++)
    for (int j = L[HY][YLO]; j <= L[HY][YHI]; j++)
      for (int k = L[HY][ZLO]; k <= L[HY][ZHI]; k++) {
        size_t id = IDX(i, j, k);
        s->Hy[id] = s->Hy[id] + s->gammaHy[id] * CURL_EY(s, id, i, k, SI);
      }
#pragma omp parallel for
  for (int i = L[HZ][XLO]; i <= L[HZ][XHI]; i++)
    for (int j = L[HZ][YLO]; j <= L[HZ][YHI]; j++)
      for (int k = L[HZ][ZLO]; k <= L[HZ][ZHI]; k++) {
        size_t id = IDX(i, j, k);
        s->Hz[id] = s->Hz[id] + s->gammaHz[id] * CURL_EZ(s, id, i, j, SI, SJ);
      }
}

/* ---- UPML slabs (pml.c:859-1176) ---- */

static size_t pidx(const int *l, int i, int j, int k)
{
  size_t nj = (size_t)(l[YHI] - l[YLO] + 1), nk = (size_t)(l[ZHI] - l[ZLO] + 1);
  return ((size_t)(i - l[XLO]) * nj + (size_t)(j - l[YLO])) * nk + (size_t)(k - l[ZLO]);
}

static void pml_e(vh_solve *s)
{
  const size_t SJ = (size_t)s->numCells[2], SI = SJ * (size_t)s->numCells[1];
  for (int r = 0; r < 6; r++) {
    if (!s->Px[r]) continue;
    int *l = s->fplim[r][EX];
#pragma omp parallel for
    for (int i = l[XLO]; i <= l[XHI]; i++)
      for (int j = l[YLO]; j <= l[YHI]; j++)
        for (int k = l[ZLO]; k <= l[ZHI]; k++) {
          size_t id = IDX(i, j, k), p = pidx(l, i, j, k);
          real oldPP = s->PPx[r][p];
          s->PPx[r][p] = s->alphaEx[id] * s->PPx[r][p] + s->betaEx[id] * CURL_HX(s, id, j, k, SJ);
          real oldP = s->Px[r][p];
          s->Px[r][p] = s->ady[j] * s->Px[r][p] + s->bdy[j] * (s->PPx[r][p] - oldPP);
          s->Ex[id] = s->adz[k] * s->Ex[id] + s->bdz[k] * s->ibhx[i] * (s->Px[r][p] - s->ahx[i] * oldP);
        }
    l = s->fplim[r][EY];
#pragma omp parallel for
    for (int i = l[XLO]; i <= l[XHI]; i++)
      for (int j = l[YLO]; j <= l[YHI]; j++)
        for (int k = l[ZLO]; k <= l[ZHI]; k++) {
          size_t id = IDX(i, j, k), p = pidx(l, i, j, k);
          real oldPP = s->PPy[r][p];
          s->PPy[r][p] = s->alphaEy[id] * s->PPy[r][p] + s->betaEy[id] * CURL_HY(s, id, i, k, SI);
          real oldP = s->Py[r][p];
          s->Py[r][p] = s->adz[k] * s->Py[r][p] + s->bdz[k] * (s->PPy[r][p] - oldPP);
          s->Ey[id] = s->adx[i] * s->Ey[id] + s->bdx[i] * s->ibhy[j] * (s->Py[r][p] - s->ahy[j] * oldP);
        }
    l = s->fplim[r][EZ];
#pragma omp parallel for
    for (int i = l[XLO]; i <= l[XHI]; i++)
      for (int j = l[YLO]; j <= l[YHI]; j++)
        for (int k = l[ZLO]; k <= l[ZHI]; k++) {
          size_t id = IDX(i, j, k), p = pidx(l, i, j, k);
          real oldPP = s->PPz[r][p];
          s->PPz[r][p] = s->alphaEz[id] * s->PPz[r][p] + s->betaEz[id] * CURL_HZ(s, id, i, j, SI, SJ);
          real oldP = s->Pz[r][p];
          s->Pz[r][p] = s->adx[i] * s->Pz[r][p] + s->bdx[i] * (s->PPz[r][p] - oldPP);
          s->Ez[id] = s->ady[j] * s->Ez[id] + s->bdy[j] * s->ibhz[k] * (s->Pz[r][p] - s->ahz[k] * oldP);
        }
  }
}

static void pml_h(vh_solve *s)
{
  const size_t SJ = (size_t)s->numCells[2], SI = SJ * (size_t)s->numCells[1];
  for (int r = 0; r < 6; r++) {
    if (!s->Bx[r]) continue;
    int *l = s->fplim[r][HX];
#pragma omp parallel for
    for (int i = l[XLO]; i <= l[XHI]; i++)
      for (int j = l[YLO]; j <= l[YHI]; j++)
        for (int k = l[ZLO]; k <= l[ZHI]; k++) {
          size_t id = IDX(i, j, k), p = pidx(l, i, j, k);
          real oldB = s->Bx[r][p];
          s->Bx[r][p] = s->ahy[j] * s->Bx[r][p] + s->gammaHx[id] * s->bhy[j] * CURL_EX(s, id, j, k, SJ);
          s->Hx[id] = s->ahz[k] * s->Hx[id] + s->bhz[k] * s->ibdx[i] * (s->Bx[r][p] - s->adx[i] * oldB);
        }
    l = s->fplim[r][HY];
#pragma omp parallel for
    for (int i = l[XLO]; i <= l[XHI]; i++)
      for (int j = l[YLO]; j <= l[YHI]; j++)
        for (int k = l[ZLO]; k <= l[ZHI]; k++) {
          size_t id = IDX(i, j, k), p = pidx(l, i, j, k);
          real oldB = s->By[r][p];
          s->By[r][p] = s->ahz[k] * s->By[r][p] + s->gammaHy[id] * s->bhz[k] * CURL_EY(s, id, i, k, SI);
          s->Hy[id] = s->ahx[i] * s->Hy[id] + s->bhx[i] * s->ibdy[j] * (s->By[r][p] - s->ady[j] * oldB);
        }
    l = s->fplim[r][HZ];
#pragma omp parallel for
    for (int i = l[XLO]; i <= l[XHI]; i++)
      for (int j = l[YLO]; j <= l[YHI]; j++)
        for (int k = l[ZLO]; k <= l[ZHI]; k++) {
          size_t id = IDX(i, j, k), p = pidx(l, i, j, k);
          real oldB = s->Bz[r][p];
          s->Bz[r][p] = s->ahx[i] * s->Bz[r][p] + s->gammaHz[id] * s->bhx[i] * CURL_EZ(s, id, i, j, SI, SJ);
          s->Hz[id] = s->ahy[j] * s->Hz[id] + s->bhy[j] * s->ibdz[k] * (s->Bz[r][p] - s->adz[k] * oldB);
        }
  }
}

/* ---- first-order Mur (mur.c:135-555) ---- */

/* Tangential E on one Mur face: E_b = E_n + zeta (E*_n - E_b), where E*_n is
 * the neighbour's E^{n+1} recomputed from H and n is the cell one step into
 * the mesh. */
static void mur_e_face(vh_solve *s, int face, int comp)
{
  const size_t SJ = (size_t)s->numCells[2], SI = SJ * (size_t)s->numCells[1];
  int *l = s->fmlim[face][comp];
  int axis = face >> 1, step = (face & 1) ? -1 : +1;
  size_t S[3] = {SI, SJ, 1};
  real *E[3] = {s->Ex, s->Ey, s->Ez}, *A[3] = {s->alphaEx, s->alphaEy, s->alphaEz};
  real *B[3] = {s->betaEx, s->betaEy, s->betaEz};
  int lo[3] = {l[XLO], l[YLO], l[ZLO]}, hi[3] = {l[XHI], l[YHI], l[ZHI]};
  lo[axis] = hi[axis] = (face & 1) ? l[2 * axis + 1] : l[2 * axis];
  for (int i = lo[0]; i <= hi[0]; i++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int k = lo[2]; k <= hi[2]; k++) {
        size_t id = IDX(i, j, k);
        size_t nb = step > 0 ? id + S[axis] : id - S[axis];
        int ni = i + (axis == 0 ? step : 0), nj = j + (axis == 1 ? step : 0), nk = k + (axis == 2 ? step : 0);
        real curl;
        if (comp == EX) curl = CURL_HX(s, nb, nj, nk, SJ);
        else if (comp == EY) curl = CURL_HY(s, nb, ni, nk, SI);
        else curl = CURL_HZ(s, nb, ni, nj, SI, SJ);
        real tmp = A[comp][nb] * E[comp][nb] + B[comp][nb] * curl;
        E[comp][id] = E[comp][nb] + s->zeta[face] * (tmp - E[comp][id]);
      }
}

static void mur_e(vh_solve *s)
{
  /* Face and component order of mur.c:144-391. */
  static const int order[6][2] = {{EY, EZ}, {EY, EZ}, {EX, EZ}, {EX, EZ}, {EY, EX}, {EY, EX}};
  for (int f = 0; f < 6; f++) {
    if (vh_outer_type(s, f) != BT_MUR) continue;
    mur_e_face(s, f, order[f][0]);
    mur_e_face(s, f, order[f][1]);
  }
}

static void mur_h(vh_solve *s)
{
  /* Normal H on each Mur face gets the ordinary update (mur.c:401-555); the
   * face box is one plane thick along its normal. */
  for (int f = 0; f < 6; f++) {
    if (vh_outer_type(s, f) != BT_MUR) continue;
    int comp = HX + (f >> 1), L[6][6];
    for (int c = 0; c < 6; c++) {
      for (int q = 0; q < 6; q++) L[c][q] = s->fmlim[f][c][q];
      if (c != comp) { L[c][XLO] = 1; L[c][XHI] = 0; }
    }
    h_box(s, (const int(*)[6])L);
  }
}

/* ---- Debye ADE (debye.c:243-307) ---- */

static void debye_e(vh_solve *s)
{
  real *E[3] = {s->Ex, s->Ey, s->Ez}, *B[3] = {s->betaEx, s->betaEy, s->betaEz};
  for (int b = 0; b < s->numDebye; b++) {
    vh_debye *d = &s->debye[b];
    vh_medium *m = &s->media[d->medium];
    int np = m->numPoles;
    for (int f = EX; f <= EZ; f++) {
      int *l = d->flim[f];
      size_t c = 0;
      for (int i = l[XLO]; i <= l[XHI]; i++)
        for (int j = l[YLO]; j <= l[YHI]; j++)
          for (int k = l[ZLO]; k <= l[ZHI]; k++, c++) {
            size_t id = IDX(i, j, k);
            double complex *J = &d->J[f][c * (size_t)np];
            double complex Jsum = 0.0 + I * 0.0;
            for (int p = 0; p < np; p++) Jsum = Jsum + (1 + m->dalpha[p]) * J[p];
            E[f][id] = E[f][id] - B[f][id] * creal(Jsum);
            for (int p = 0; p < np; p++) J[p] = m->dalpha[p] * J[p] + m->dbeta[p] * (E[f][id] - d->lastE[f][c]);
            d->lastE[f][c] = E[f][id];
          }
    }
  }
}

/* ---- SIBC (sibc.c:443-776, filter.c:309-327) ---- */

static const real Acp[4][4] = {{0, 0, 1, 0}, {0, 0, 0, -1}, {-1, 0, 0, 0}, {0, 1, 0, 0}};
static const real Acm[4][4] = {{0, 0, 0, -1}, {0, 0, 1, 0}, {0, 1, 0, 0}, {-1, 0, 0, 0}};
static const real Asp[4][4] = {{-1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, -1, 0}, {0, 0, 0, 1}};
static const real Asm[4][4] = {{0, 1, 0, 0}, {-1, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, -1, 0}};
static const real Bcp[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
static const real Bcm[4][4] = {{0, 1, 0, 0}, {1, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 1, 0}};
static const real Bsp[4][4] = {{0, 0, -1, 0}, {0, 0, 0, -1}, {1, 0, 0, 0}, {0, 1, 0, 0}};
static const real Bsm[4][4] = {{0, 0, 0, -1}, {0, 0, -1, 0}, {0, 1, 0, 0}, {1, 0, 0, 0}};

static void mat_vec(real y[4], real A[4][4], const real x[4])
{
  for (int p = 0; p < 4; p++) {
    y[p] = 0.0;
    for (int q = 0; q < 4; q++) y[p] = y[p] + A[p][q] * x[q];
  }
}

static void mat_comb(real C[4][4], real a, const real A[4][4], real b, const real B[4][4])
{
  for (int p = 0; p < 4; p++)
    for (int q = 0; q < 4; q++) C[p][q] = a * A[p][q] + b * B[p][q];
}

static double rc_step(const vh_rcfilter *rc, double complex *zeta, double *old, double x)
{
  double y = rc->asymp * x;
  for (int k = 0; k < rc->numPoles; k++) {
    zeta[k] = rc->q0[k] * zeta[k] + rc->q1[k] * *old + rc->q2[k] * x;
    y = y + zeta[k];
  }
  *old = x;
  return y;
}

static void sibc_e(vh_solve *s)
{
  for (int n = 0; n < s->numSibc; n++) {
    vh_sibc *sb = &s->sibc[n];
    vh_boundary *bd = &s->bounds[sb->boundary];
    real A[4][4], B[4][4];
    if (sb->orientation == 1) {
      mat_comb(A, sb->cosa, Acp, sb->sina, Asp);
      mat_comb(B, sb->cosa, Bcp, sb->sina, Bsp);
    } else {
      mat_comb(A, sb->cosa, Acm, sb->sina, Asm);
      mat_comb(B, sb->cosa, Bcm, sb->sina, Bsm);
    }
    size_t f = 0;
    for (int i = sb->gbbox[XLO]; i < sb->gbbox[XHI]; i++)
      for (int j = sb->gbbox[YLO]; j < sb->gbbox[YHI]; j++)
        for (int k = sb->gbbox[ZLO]; k < sb->gbbox[ZHI]; k++, f++) {
          bool *a = &sb->isAdjA[4 * f], *b = &sb->isAdjB[4 * f];
          real Htan[4], Hin[4], Eout[4];
          switch (sb->normal) {
          case XDIR:
            s->Ey[IDX(i, j, k)] = 0.0; s->Ey[IDX(i, j, k + 1)] = 0.0;
            s->Ez[IDX(i, j, k)] = 0.0; s->Ez[IDX(i, j + 1, k)] = 0.0;
            Htan[0] = 0.5 * (1 + a[0] + a[1]) * (s->Hy[IDX(i - 1, j, k)] + s->Hy[IDX(i - 1, j + 1, k)]);
            Htan[1] = 0.5 * (1 + b[0] + b[1]) * (s->Hy[IDX(i, j, k)] + s->Hy[IDX(i, j + 1, k)]);
            Htan[2] = 0.5 * (1 + a[2] + a[3]) * (s->Hz[IDX(i - 1, j, k)] + s->Hz[IDX(i - 1, j, k + 1)]);
            Htan[3] = 0.5 * (1 + b[2] + b[3]) * (s->Hz[IDX(i, j, k)] + s->Hz[IDX(i, j, k + 1)]);
            break;
          case YDIR:
            s->Ez[IDX(i, j, k)] = 0.0; s->Ez[IDX(i + 1, j, k)] = 0.0;
            s->Ex[IDX(i, j, k)] = 0.0; s->Ex[IDX(i, j, k + 1)] = 0.0;
            Htan[0] = 0.5 * (1 + a[0] + a[1]) * (s->Hz[IDX(i, j - 1, k)] + s->Hz[IDX(i, j - 1, k + 1)]);
            Htan[1] = 0.5 * (1 + b[0] + b[1]) * (s->Hz[IDX(i, j, k)] + s->Hz[IDX(i, j, k + 1)]);
            Htan[2] = 0.5 * (1 + a[2] + a[3]) * (s->Hx[IDX(i, j - 1, k)] + s->Hx[IDX(i + 1, j - 1, k)]);
            Htan[3] = 0.5 * (1 + b[2] + b[3]) * (s->Hx[IDX(i, j, k)] + s->Hx[IDX(i + 1, j, k)]);
            break;
          default:
            s->Ex[IDX(i, j, k)] = 0.0; s->Ex[IDX(i, j + 1, k)] = 0.0;
            s->Ey[IDX(i, j, k)] = 0.0; s->Ey[IDX(i + 1, j, k)] = 0.0;
            Htan[0] = 0.5 * (1 + a[0] + a[1]) * (s->Hx[IDX(i, j, k - 1)] + s->Hx[IDX(i + 1, j, k - 1)]);
            Htan[1] = 0.5 * (1 + b[0] + b[1]) * (s->Hx[IDX(i, j, k)] + s->Hx[IDX(i + 1, j, k)]);
            Htan[2] = 0.5 * (1 + a[2] + a[3]) * (s->Hy[IDX(i, j, k - 1)] + s->Hy[IDX(i, j + 1, k - 1)]);
            Htan[3] = 0.5 * (1 + b[2] + b[3]) * (s->Hy[IDX(i, j, k)] + s->Hy[IDX(i, j + 1, k)]);
            break;
          }
          mat_vec(Hin, A, Htan);
          for (int p = 0; p < 4; p++) {
            Eout[p] = 0.0;
            for (int q = 0; q < 4; q++) {
              size_t fi = f * 16 + (size_t)(4 * p + q);
              Eout[p] = Eout[p] + rc_step(&bd->rc[4 * p + q], &sb->zeta[fi * (size_t)sb->maxPoles], &sb->old[fi], Hin[q]);
            }
          }
          mat_vec(&sb->Etan[4 * f], B, Eout);
        }
  }
}

static void sibc_h(vh_solve *s)
{
  for (int n = 0; n < s->numSibc; n++) {
    vh_sibc *sb = &s->sibc[n];
    size_t f = 0;
    for (int i = sb->gbbox[XLO]; i < sb->gbbox[XHI]; i++)
      for (int j = sb->gbbox[YLO]; j < sb->gbbox[YHI]; j++)
        for (int k = sb->gbbox[ZLO]; k < sb->gbbox[ZHI]; k++, f++) {
          bool *A = &sb->isAdjA[4 * f], *B = &sb->isAdjB[4 * f];
          real *Et = &sb->Etan[4 * f];
          real wxl, wxh, wyl, wyh, wzl, wzh;
          switch (sb->normal) {
          case XDIR: {
            s->Hx[IDX(i, j, k)] = 0.0;
            wyl = 0.5 * (1 + vh_is_pmc_edge(s, ZDIR, k));
            wyh = 0.5 * (1 + vh_is_pmc_edge(s, ZDIR, k + 1));
            wzl = 0.5 * (1 + vh_is_pmc_edge(s, YDIR, j));
            wzh = 0.5 * (1 + vh_is_pmc_edge(s, YDIR, j + 1));
            real Eya = Et[0], Eyb = Et[1], Eza = Et[2], Ezb = Et[3];
            s->Hy[IDX(i - 1, j, k)] = s->Hy[IDX(i - 1, j, k)] + s->gammaHy[IDX(i - 1, j, k)] * wzl * (s->idex[i - 1] * Eza);
            s->Hy[IDX(i - 1, j + 1, k)] = s->Hy[IDX(i - 1, j + 1, k)] + s->gammaHy[IDX(i - 1, j + 1, k)] * wzh * (s->idex[i - 1] * Eza);
            s->Hy[IDX(i, j, k)] = s->Hy[IDX(i, j, k)] - s->gammaHy[IDX(i, j, k)] * wzl * (s->idex[i] * Ezb);
            s->Hy[IDX(i, j + 1, k)] = s->Hy[IDX(i, j + 1, k)] - s->gammaHy[IDX(i, j + 1, k)] * wzh * (s->idex[i] * Ezb);
            s->Hz[IDX(i - 1, j, k)] = s->Hz[IDX(i - 1, j, k)] - s->gammaHz[IDX(i - 1, j, k)] * wyl * (s->idex[i - 1] * Eya);
            s->Hz[IDX(i - 1, j, k + 1)] = s->Hz[IDX(i - 1, j, k + 1)] - s->gammaHz[IDX(i - 1, j, k + 1)] * wyh * (s->idex[i - 1] * Eya);
            s->Hz[IDX(i, j, k)] = s->Hz[IDX(i, j, k)] + s->gammaHz[IDX(i, j, k)] * wyl * (s->idex[i] * Eyb);
            s->Hz[IDX(i, j, k + 1)] = s->Hz[IDX(i, j, k + 1)] + s->gammaHz[IDX(i, j, k + 1)] * wyh * (s->idex[i] * Eyb);
            s->Hx[IDX(i, j + 1, k)] = s->Hx[IDX(i, j + 1, k)] - s->gammaHx[IDX(i, j + 1, k)] * 0.5 * wzh * (!A[0] * (s->idex[i - 1] * Eza) + !B[0] * (s->idex[i] * Ezb));
            s->Hx[IDX(i, j - 1, k)] = s->Hx[IDX(i, j - 1, k)] + s->gammaHx[IDX(i, j - 1, k)] * 0.5 * wzl * (!A[1] * (s->idex[i - 1] * Eza) + !B[1] * (s->idex[i] * Ezb));
            s->Hx[IDX(i, j, k + 1)] = s->Hx[IDX(i, j, k + 1)] - s->gammaHx[IDX(i, j, k + 1)] * 0.5 * wyh * (!A[2] * (s->idex[i - 1] * Eya) + !B[2] * (s->idex[i] * Eyb));
            s->Hx[IDX(i, j, k - 1)] = s->Hx[IDX(i, j, k - 1)] + s->gammaHx[IDX(i, j, k - 1)] * 0.5 * wyl * (!A[3] * (s->idex[i - 1] * Eya) + !B[3] * (s->idex[i] * Eyb));
            break;
          }
          case YDIR: {
            s->Hy[IDX(i, j, k)] = 0.0;
            wzl = 0.5 * (1 + vh_is_pmc_edge(s, XDIR, i));
            wzh = 0.5 * (1 + vh_is_pmc_edge(s, XDIR, i + 1));
            wxl = 0.5 * (1 + vh_is_pmc_edge(s, ZDIR, k));
            wxh = 0.5 * (1 + vh_is_pmc_edge(s, ZDIR, k + 1));
            real Eza = Et[0], Ezb = Et[1], Exa = Et[2], Exb = Et[3];
            s->Hz[IDX(i, j - 1, k)] = s->Hz[IDX(i, j - 1, k)] + s->gammaHz[IDX(i, j - 1, k)] * wxl * (s->idey[j - 1] * Exa);
            s->Hz[IDX(i, j - 1, k + 1)] = s->Hz[IDX(i, j - 1, k + 1)] + s->gammaHz[IDX(i, j - 1, k + 1)] * wxh * (s->idey[j - 1] * Exa);
            s->Hz[IDX(i, j, k)] = s->Hz[IDX(i, j, k)] - s->gammaHz[IDX(i, j, k)] * wxl * (s->idey[j] * Exb);
            s->Hz[IDX(i, j, k + 1)] = s->Hz[IDX(i, j, k + 1)] - s->gammaHz[IDX(i, j, k + 1)] * wxh * (s->idey[j] * Exb);
            s->Hx[IDX(i, j - 1, k)] = s->Hx[IDX(i, j - 1, k)] - s->gammaHx[IDX(i, j - 1, k)] * wzl * (s->idey[j - 1] * Eza);
            s->Hx[IDX(i + 1, j - 1, k)] = s->Hx[IDX(i + 1, j - 1, k)] - s->gammaHx[IDX(i + 1, j - 1, k)] * wzh * (s->idey[j - 1] * Eza);
            s->Hx[IDX(i, j, k)] = s->Hx[IDX(i, j, k)] + s->gammaHx[IDX(i, j, k)] * wzl * (s->idey[j] * Ezb);
            s->Hx[IDX(i + 1, j, k)] = s->Hx[IDX(i + 1, j, k)] + s->gammaHx[IDX(i + 1, j, k)] * wzh * (s->idey[j] * Ezb);
            /* The B-side term of the first two uses Exb (sibc.c:689-690). */
            s->Hy[IDX(i, j, k + 1)] = s->Hy[IDX(i, j, k + 1)] + s->gammaHy[IDX(i, j, k + 1)] * 0.5 * wzh * (!A[0] * (s->idey[j - 1] * Eza) + !B[0] * (s->idey[j] * Exb));
            s->Hy[IDX(i, j, k - 1)] = s->Hy[IDX(i, j, k - 1)] - s->gammaHy[IDX(i, j, k - 1)] * 0.5 * wzl * (!A[1] * (s->idey[j - 1] * Eza) + !B[1] * (s->idey[j] * Exb));
            s->Hy[IDX(i + 1, j, k)] = s->Hy[IDX(i + 1, j, k)] + s->gammaHy[IDX(i + 1, j, k)] * 0.5 * wxh * (!A[2] * (s->idey[j - 1] * Exa) + !B[2] * (s->idey[j] * Exb));
            s->Hy[IDX(i - 1, j, k)] = s->Hy[IDX(i - 1, j, k)] - s->gammaHy[IDX(i - 1, j, k)] * 0.5 * wxl * (!A[3] * (s->idey[j - 1] * Exa) + !B[3] * (s->idey[j] * Exb));
            break;
          }
          default: {
            s->Hz[IDX(i, j, k)] = 0.0;
            wxl = 0.5 * (1 + vh_is_pmc_edge(s, YDIR, j));
            wxh = 0.5 * (1 + vh_is_pmc_edge(s, YDIR, j + 1));
            wyl = 0.5 * (1 + vh_is_pmc_edge(s, XDIR, i));
            wyh = 0.5 * (1 + vh_is_pmc_edge(s, XDIR, i + 1));
            real Exa = Et[0], Exb = Et[1], Eya = Et[2], Eyb = Et[3];
            s->Hx[IDX(i, j, k - 1)] = s->Hx[IDX(i, j, k - 1)] + s->gammaHx[IDX(i, j, k - 1)] * wyl * (s->idez[k - 1] * Eya);
            s->Hx[IDX(i + 1, j, k - 1)] = s->Hx[IDX(i + 1, j, k - 1)] + s->gammaHx[IDX(i + 1, j, k - 1)] * wyh * (s->idez[k - 1] * Eya);
            s->Hx[IDX(i, j, k)] = s->Hx[IDX(i, j, k)] - s->gammaHx[IDX(i, j, k)] * wyl * (s->idez[k] * Eyb);
            s->Hx[IDX(i + 1, j, k)] = s->Hx[IDX(i + 1, j, k)] - s->gammaHx[IDX(i + 1, j, k)] * wyh * (s->idez[k] * Eyb);
            s->Hy[IDX(i, j, k - 1)] = s->Hy[IDX(i, j, k - 1)] - s->gammaHy[IDX(i, j, k - 1)] * wxl * (s->idez[k - 1] * Exa);
            s->Hy[IDX(i, j + 1, k - 1)] = s->Hy[IDX(i, j + 1, k - 1)] - s->gammaHy[IDX(i, j + 1, k - 1)] * wxh * (s->idez[k - 1] * Exa);
            s->Hy[IDX(i, j, k)] = s->Hy[IDX(i, j, k)] + s->gammaHy[IDX(i, j, k)] * wxl * (s->idez[k] * Exb);
            s->Hy[IDX(i, j + 1, k)] = s->Hy[IDX(i, j + 1, k)] + s->gammaHy[IDX(i, j + 1, k)] * wxh * (s->idez[k] * Exb);
            s->Hz[IDX(i + 1, j, k)] = s->Hz[IDX(i + 1, j, k)] + s->gammaHz[IDX(i + 1, j, k)] * 0.5 * wyh * (!A[0] * (s->idez[k - 1] * Eya) + !B[0] * (s->idez[k] * Eyb));
            s->Hz[IDX(i - 1, j, k)] = s->Hz[IDX(i - 1, j, k)] - s->gammaHz[IDX(i - 1, j, k)] * 0.5 * wyl * (!A[1] * (s->idez[k - 1] * Eya) + !B[1] * (s->idez[k] * Eyb));
            s->Hz[IDX(i, j + 1, k)] = s->Hz[IDX(i, j + 1, k)] + s->gammaHz[IDX(i, j + 1, k)] * 0.5 * wxh * (!A[2] * (s->idez[k - 1] * Exa) + !B[2] * (s->idez[k] * Exb));
            s->Hz[IDX(i, j - 1, k)] = s->Hz[IDX(i, j - 1, k)] - s->gammaHz[IDX(i, j - 1, k)] * 0.5 * wxl * (!A[3] * (s->idez[k - 1] * Exa) + !B[3] * (s->idez[k] * Exb));
            break;
          }
          }
        }
  }
  for (int n = 0; n < s->numSibc; n++) {
    vh_sibc *sb = &s->sibc[n];
    real *Hn = sb->normal == XDIR ? s->Hx : (sb->normal == YDIR ? s->Hy : s->Hz);
    for (int i = sb->gbbox[XLO]; i < sb->gbbox[XHI]; i++)
      for (int j = sb->gbbox[YLO]; j < sb->gbbox[YHI]; j++)
        for (int k = sb->gbbox[ZLO]; k < sb->gbbox[ZHI]; k++) Hn[IDX(i, j, k)] = 0.0;
  }
}

/* ---- sources (source.c:602-753) ---- */

static void sources(vh_solve *s, real t, int electric)
{
  real *F[6] = {s->Ex, s->Ey, s->Ez, s->Hx, s->Hy, s->Hz};
  real *C[6] = {s->betaEx, s->betaEy, s->betaEz, s->gammaHx, s->gammaHy, s->gammaHz};
  for (int n = 0; n < s->numSources; n++) {
    vh_source *it = &s->sources[n];
    real source = it->size * vh_waveform_value(s, t, it->waveform, it->delay);
    real isSoft = (real)it->isSoft;
    int field = it->field;
    int plainType = electric ? ST_EFIELD : ST_HFIELD;
    int densType = electric ? ST_ELEC_CURR_DENSITY : ST_MAGN_CURR_DENSITY;
    if (it->type != plainType && it->type != densType) continue;
    if ((field < HX) != (electric != 0)) continue;
    int *l = it->flim[field];
    for (int i = l[XLO]; i <= l[XHI]; i++)
      for (int j = l[YLO]; j <= l[YHI]; j++)
        for (int k = l[ZLO]; k <= l[ZHI]; k++) {
          size_t id = IDX(i, j, k);
          if (it->type == plainType) F[field][id] = isSoft * F[field][id] + source;
          else F[field][id] = isSoft * F[field][id] - C[field][id] * source;
        }
  }
}

/* ---- TF/SF plane waves (planewave.c:557-1007, 1315-1419) ---- */

static void aux_e(vh_solve *s, vh_planewave *pw, real t)
{
  for (int i = 1; i < pw->xb; i++) pw->Eyi[i] = pw->Eyi[i] + pw->betaEyi * (pw->Hzi[i - 1] - pw->Hzi[i]);
  for (int i = pw->xb; i < pw->nx; i++) {
    int Lp = i - pw->xb;
    real oldPP = pw->PPyi[Lp];
    pw->PPyi[Lp] = pw->PPyi[Lp] + pw->betaEyi * (pw->Hzi[i - 1] - pw->Hzi[i]);
    real oldP = pw->Pyi[Lp];
    pw->Pyi[Lp] = pw->Pyi[Lp] + (pw->PPyi[Lp] - oldPP);
    pw->Eyi[i] = pw->adx[Lp] * pw->Eyi[i] + pw->bdx[Lp] * (pw->Pyi[Lp] - oldP);
  }
  real w = vh_waveform_value(s, t, pw->waveform, pw->delay);
  pw->Eyi[0] = w; /* m0 - 2 with m0 = 2 */
}

static void aux_h(vh_solve *s, vh_planewave *pw)
{
  (void)s;
  for (int i = 0; i < pw->xb; i++) pw->Hzi[i] = pw->Hzi[i] + pw->gammaHzi * (pw->Eyi[i] - pw->Eyi[i + 1]);
  for (int i = pw->xb; i < pw->nx; i++) {
    int Lp = i - pw->xb;
    real oldB = pw->Bzi[Lp];
    pw->Bzi[Lp] = pw->ahx[Lp] * pw->Bzi[Lp] + pw->gammaHzi * pw->bhx[Lp] * (-pw->Eyi[i + 1] + pw->Eyi[i]);
    pw->Hzi[i] = pw->Hzi[i] + (pw->Bzi[Lp] - oldB);
  }
}

static real incident(vh_solve *s, vh_planewave *pw, int field, int i, int j, int k, real t)
{
  real r[3], d;
  if (s->useAuxGrid) {
    vh_field_index_location(r, field, i, j, k);
    d = pw->kinc[0] * (r[0] - pw->ijk0[0]) + pw->kinc[1] * (r[1] - pw->ijk0[1]) + pw->kinc[2] * (r[2] - pw->ijk0[2]);
    real value, dp;
    int id;
    const int m0 = 2;
    if (field < HX) {
      id = floor(d);
      dp = d - id;
      value = (1 - dp) * pw->Eyi[m0 + id] + dp * pw->Eyi[m0 + id + 1];
    } else {
      id = floor(d + 0.5);
      dp = d + 0.5 - id;
      value = s->eta0 * ((1 - dp) * pw->Hzi[m0 - 1 + id] + dp * pw->Hzi[m0 + id]);
    }
    return pw->Finc[field] * value;
  }
  vh_field_physical_location(s, r, field, i, j, k);
  d = pw->kinc[0] * (r[0] - pw->r0[0]) + pw->kinc[1] * (r[1] - pw->r0[1]) + pw->kinc[2] * (r[2] - pw->r0[2]);
  return pw->Finc[field] * vh_waveform_value(s, t - d / pw->phaseVelocity, pw->waveform, pw->delay);
}

/*
 * One TF/SF face/component correction.  E corrections multiply by alpha
 * again: E = alpha*E + sign*beta*(1/dh * Finc) (planewave.c:601 etc.).
 * `axis` is the face normal, `inc` the incident component, `off` the offset
 * of the incident sample along the normal, `sign` the correction sign.
 */
static void tfsf_e_face(vh_solve *s, vh_planewave *pw, int face, int comp, int inc, int off, int sign, real t)
{
  int *l = pw->flim[face][comp];
  int axis = face >> 1;
  int plane = (face & 1) ? l[2 * axis + 1] : l[2 * axis];
  real *E[3] = {s->Ex, s->Ey, s->Ez}, *A[3] = {s->alphaEx, s->alphaEy, s->alphaEz};
  real *B[3] = {s->betaEx, s->betaEy, s->betaEz};
  const real *idh[3] = {s->idhx, s->idhy, s->idhz};
  int lo[3] = {l[XLO], l[YLO], l[ZLO]}, hi[3] = {l[XHI], l[YHI], l[ZHI]};
  lo[axis] = hi[axis] = plane;
  for (int i = lo[0]; i <= hi[0]; i++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int k = lo[2]; k <= hi[2]; k++) {
        int ii = i + (axis == 0 ? off : 0), jj = j + (axis == 1 ? off : 0), kk = k + (axis == 2 ? off : 0);
        real incField = incident(s, pw, inc, ii, jj, kk, t);
        int n3[3] = {i, j, k};
        size_t id = IDX(i, j, k);
        if (sign > 0) E[comp][id] = A[comp][id] * E[comp][id] + B[comp][id] * (idh[axis][n3[axis]] * incField);
        else E[comp][id] = A[comp][id] * E[comp][id] - B[comp][id] * (idh[axis][n3[axis]] * incField);
      }
}

static void tfsf_h_face(vh_solve *s, vh_planewave *pw, int face, int comp, int inc, int off, int sign, real t)
{
  int *l = pw->flim[face][comp];
  int axis = face >> 1;
  int plane = (face & 1) ? l[2 * axis + 1] : l[2 * axis];
  real *H[3] = {s->Hx, s->Hy, s->Hz}, *G[3] = {s->gammaHx, s->gammaHy, s->gammaHz};
  const real *ide[3] = {s->idex, s->idey, s->idez};
  int lo[3] = {l[XLO], l[YLO], l[ZLO]}, hi[3] = {l[XHI], l[YHI], l[ZHI]};
  lo[axis] = hi[axis] = plane;
  int c = comp - HX;
  for (int i = lo[0]; i <= hi[0]; i++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int k = lo[2]; k <= hi[2]; k++) {
        int ii = i + (axis == 0 ? off : 0), jj = j + (axis == 1 ? off : 0), kk = k + (axis == 2 ? off : 0);
        real incField = incident(s, pw, inc, ii, jj, kk, t);
        int n3[3] = {i, j, k};
        size_t id = IDX(i, j, k);
        if (sign > 0) H[c][id] = H[c][id] + G[c][id] * (ide[axis][n3[axis]] * incField);
        else H[c][id] = H[c][id] - G[c][id] * (ide[axis][n3[axis]] * incField);
      }
}

/* Face order YLO, YHI, ZLO, ZHI, XLO, XHI and the signs of eqns (5.48)-(5.59). */
static const struct { int face, comp, inc, off, sign; } TFSF_E[12] = {
  {YLO, EX, HZ, -1, -1}, {YLO, EZ, HX, -1, +1}, {YHI, EX, HZ, 0, +1}, {YHI, EZ, HX, 0, -1},
  {ZLO, EX, HY, -1, +1}, {ZLO, EY, HX, -1, -1}, {ZHI, EX, HY, 0, -1}, {ZHI, EY, HX, 0, +1},
  {XLO, EY, HZ, -1, +1}, {XLO, EZ, HY, -1, -1}, {XHI, EY, HZ, 0, -1}, {XHI, EZ, HY, 0, +1}};
static const struct { int face, comp, inc, off, sign; } TFSF_H[12] = {
  {YLO, HZ, EX, +1, -1}, {YLO, HX, EZ, +1, +1}, {YHI, HZ, EX, 0, +1}, {YHI, HX, EZ, 0, -1},
  {ZLO, HY, EX, +1, +1}, {ZLO, HX, EY, +1, -1}, {ZHI, HY, EX, 0, -1}, {ZHI, HX, EY, 0, +1},
  {XLO, HZ, EY, +1, +1}, {XLO, HY, EZ, +1, -1}, {XHI, HZ, EY, 0, -1}, {XHI, HY, EZ, 0, +1}};

static void planewaves_e(vh_solve *s, real t)
{
  for (int p = 0; p < s->numPws; p++) {
    vh_planewave *pw = &s->pws[p];
    if (s->useAuxGrid) aux_e(s, pw, t);
    for (int q = 0; q < 12; q++)
      if (pw->isActive[TFSF_E[q].face])
        tfsf_e_face(s, pw, TFSF_E[q].face, TFSF_E[q].comp, TFSF_E[q].inc, TFSF_E[q].off, TFSF_E[q].sign, t);
  }
}

static void planewaves_h(vh_solve *s, real t)
{
  for (int p = 0; p < s->numPws; p++) {
    vh_planewave *pw = &s->pws[p];
    if (s->useAuxGrid) aux_h(s, pw);
    for (int q = 0; q < 12; q++)
      if (pw->isActive[TFSF_H[q].face])
        tfsf_h_face(s, pw, TFSF_H[q].face, TFSF_H[q].comp, TFSF_H[q].inc, TFSF_H[q].off, TFSF_H[q].sign, t);
  }
}

/* ---- ghost cells for PEC/PMC/periodic faces (surface.c:569-812) ---- */

static void ghost_plane(vh_solve *s, real *F, int face, int srcPlane, real sign)
{
  int a = face >> 1;
  int dst = (face & 1) ? s->gobox[face] : s->gobox[face] - 1;
  int lo[3] = {s->gobox[XLO] - 1, s->gobox[YLO] - 1, s->gobox[ZLO] - 1};
  int hi[3] = {s->gobox[XHI], s->gobox[YHI], s->gobox[ZHI]};
  lo[a] = hi[a] = 0;
  for (int u = lo[0]; u <= hi[0]; u++)
    for (int v = lo[1]; v <= hi[1]; v++)
      for (int w = lo[2]; w <= hi[2]; w++) {
        int d3[3] = {u, v, w}, s3[3] = {u, v, w};
        d3[a] = dst;
        s3[a] = srcPlane;
        F[IDX(d3[0], d3[1], d3[2])] = sign * F[IDX(s3[0], s3[1], s3[2])];
      }
}

static void ghost_e(vh_solve *s)
{
  real *E[3] = {s->Ex, s->Ey, s->Ez};
  for (int f = 0; f < 6; f++) {
    int a = f >> 1, t = vh_outer_type(s, f);
    int inner = (f & 1) ? s->gobox[f] - 1 : s->gobox[f];
    int opp = (f & 1) ? s->gobox[f - 1] : s->gobox[f + 1] - 1;
    if (t == BT_PEC) ghost_plane(s, E[a], f, inner, 1.0);
    else if (t == BT_PMC) ghost_plane(s, E[a], f, inner, -1.0);
    else if (t == BT_PERIODIC) ghost_plane(s, E[a], f, opp, 1.0);
  }
}

static void ghost_h(vh_solve *s)
{
  real *H[3] = {s->Hx, s->Hy, s->Hz};
  for (int f = 0; f < 6; f++) {
    int a = f >> 1, t = vh_outer_type(s, f);
    int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
    /* tangential H pairs in the reference's order: x:(Hy,Hz) y:(Hz,Hx) z:(Hx,Hy) */
    int inner = (f & 1) ? s->gobox[f] - 1 : s->gobox[f];
    int opp = (f & 1) ? s->gobox[f - 1] : s->gobox[f + 1] - 1;
    if (t == BT_PMC) { ghost_plane(s, H[b1], f, inner, -1.0); ghost_plane(s, H[b2], f, inner, -1.0); }
    else if (t == BT_PERIODIC) { ghost_plane(s, H[b1], f, opp, 1.0); ghost_plane(s, H[b2], f, opp, 1.0); }
    else if (t == BT_PEC) { ghost_plane(s, H[b1], f, inner, 1.0); ghost_plane(s, H[b2], f, inner, 1.0); }
  }
}

/* ---- time loop (simulation.c:136-230) ---- */

void cpu_ref_step(vh_solve *s, unsigned long n)
{
  real dt = s->dt;
  real timeE = n * dt;
  real timeH = (n + 0.5) * dt;
  vh_observe_step(s, n, timeE, NULL, NULL);
  pml_e(s);
  mur_e(s);
  grid_e(s);
  debye_e(s);
  sibc_e(s);
  sources(s, timeE, 1);
  planewaves_e(s, timeE);
  ghost_e(s);
  pml_h(s);
  mur_h(s);
  h_box(s, (const int(*)[6])s->gfilim);
  sibc_h(s);
  sources(s, timeH, 0);
  planewaves_h(s, timeH);
  ghost_h(s);
}

void cpu_ref_propagate(vh_solve *s, unsigned long first, unsigned long count)
{
  for (unsigned long n = first; n < first + count; n++) cpu_ref_step(s, n);
}
