/*
 * Solver set-up: everything the reference does once between reading the mesh
 * and entering propagate() (vulture.c:100-138), restated over one vh_solve.
 *
 * Expressions keep the reference's operand types and evaluation order
 * (double literals promote `real` operands exactly where the reference's do),
 * so a float build of this file reproduces the reference float build's
 * coefficients bit for bit and a double build gives the FP64 values.
 */
#include <sys/stat.h>
#include "vh_internal.h"

/* ---------------------------------------------------------------------- */
/* Bounding-box helpers (bounding_box.c:37-213).                          */
/* ---------------------------------------------------------------------- */

int vh_bbox_type(const int b[6])
{
  static const int type[8] = {BB_VOLUME, BB_SURFACE, BB_SURFACE, BB_LINE,
                              BB_SURFACE, BB_LINE, BB_LINE, BB_POINT};
  int m = (b[XLO] == b[XHI]) + 2 * (b[YLO] == b[YHI]) + 4 * (b[ZLO] == b[ZHI]);
  return type[m];
}

int vh_bbox_direction(const int b[6])
{
  static const int dir[8] = {CA_UNDEFINED, XDIR, YDIR, ZDIR, ZDIR, YDIR, XDIR, CA_UNDEFINED};
  int m = (b[XLO] == b[XHI]) + 2 * (b[YLO] == b[YHI]) + 4 * (b[ZLO] == b[ZHI]);
  return dir[m];
}

void vh_face_of_bbox(int fb[6], const int b[6], int face)
{
  for (int f = 0; f < 6; f++) fb[f] = b[f];
  int lo = face & ~1, hi = lo + 1;
  fb[lo] = fb[hi] = b[face];
}

void vh_offset_bbox(int out[6], const int b[6], const int off[6])
{
  out[XLO] = b[XLO] + off[XLO]; out[XHI] = b[XHI] + off[XLO];
  out[YLO] = b[YLO] + off[YLO]; out[YHI] = b[YHI] + off[YLO];
  out[ZLO] = b[ZLO] + off[ZLO]; out[ZHI] = b[ZHI] + off[ZLO];
}

/* A field component lies in a face when it is tangential E or normal H. */
bool vh_field_is_in_boundary(int field, int face)
{
  int axis = face >> 1;
  return field < HX ? (field != axis) : (field - HX == axis);
}

void vh_face_mask_to_bool(bool inc[6], FaceMask mask)
{
  static const FaceMask bit[6] = {0x20, 0x10, 0x8, 0x4, 0x2, 0x1};
  for (int f = 0; f < 6; f++) inc[f] = (mask & bit[f]) != 0;
}

/* grid.c:414-454 */
void vh_set_field_limits(const int b[6], int lim[6][6], const bool inc[6])
{
  for (int field = EX; field <= HZ; field++)
    for (int face = XLO; face <= ZHI; face++) {
      bool inb = vh_field_is_in_boundary(field, face);
      if ((face & 1) == 0)
        lim[field][face] = (inb && !inc[face]) ? b[face] + 1 : b[face];
      else
        lim[field][face] = (inb && inc[face]) ? b[face] : b[face] - 1;
    }
}

int vh_outer_type(const vh_solve *s, int face) { return s->bounds[s->extBoundary[face]].type; }

static int outer_layers(const vh_solve *s, int face) { return s->bounds[s->extBoundary[face]].numLayers; }

/* surface.c:1022-1052 */
bool vh_is_pmc_edge(const vh_solve *s, int dir, int index)
{
  int lo = 2 * dir, hi = lo + 1;
  int tl = vh_outer_type(s, lo), th = vh_outer_type(s, hi);
  return ((tl == BT_PMC || tl == BT_PERIODIC) && index == s->gibox[lo]) ||
         ((th == BT_PMC || th == BT_PERIODIC) && index == s->gibox[hi]);
}

/* ---------------------------------------------------------------------- */
/* Grid (grid.c:327-679).                                                  */
/* ---------------------------------------------------------------------- */

static void init_external_surface_parameters(vh_solve *s)
{
  static const char *FACE_NAME[6] = {"XLO", "XHI", "YLO", "YHI", "ZLO", "ZHI"};
  for (int f = 0; f < 6; f++) {
    int b = vh_find_boundary(s, FACE_NAME[f]);
    if (b < 0) {
      int numLayers, order;
      real n_eff, refCoeff, kmax;
      numLayers = 6; order = 4; n_eff = 1.0; refCoeff = -1.0; kmax = 1.0;
      b = vh_add_boundary(s, FACE_NAME[f], BT_PML, numLayers, order, n_eff, refCoeff, kmax, "", NULL, NULL);
    }
    s->extBoundary[f] = b;
  }
  for (int f = 0; f < 6; f++) {
    vh_boundary *b = &s->bounds[s->extBoundary[f]];
    switch (b->type) {
    case BT_PEC: case BT_PMC: case BT_MUR: b->numLayers = 0; break;
    case BT_PERIODIC: {
      b->numLayers = 0;
      int opp = f - 2 * (f % 2) + 1;
      if (vh_outer_type(s, opp) != BT_PERIODIC)
        vh_fail(s, "PERIODIC boundary on face %d does not match the opposite face", f);
      break;
    }
    case BT_PML:
      if (b->numLayers < 1) vh_fail(s, "PML on face %d has fewer than one layer", f);
      break;
    default: vh_fail(s, "unsupported external boundary type %d on face %d", b->type, f);
    }
  }
}

static void set_grid_extents(vh_solve *s)
{
  for (int a = 0; a < 3; a++) {
    int lo = 2 * a, hi = lo + 1;
    s->ggbox[lo] = 0;
    s->gobox[lo] = 1;
    s->gibox[lo] = s->gobox[lo] + outer_layers(s, lo);
    s->gibox[hi] = s->gibox[lo] + s->numLines[a] - 1;
    s->gobox[hi] = s->gibox[hi] + outer_layers(s, hi);
    s->ggbox[hi] = s->gobox[hi] + 1;
    s->numCells[a] = s->gobox[hi] - s->gobox[lo] + 2;
  }
  s->ncell = (size_t)s->numCells[0] * s->numCells[1] * s->numCells[2];
}

static void init_field_array_limits(vh_solve *s)
{
  bool incInner[6], incOuter[6], incGhost[6];
  for (int f = 0; f < 6; f++) {
    incInner[f] = vh_outer_type(s, f) != BT_MUR;
    incOuter[f] = true;
    incGhost[f] = false;
  }
  vh_set_field_limits(s->gibox, s->gfilim, incInner);
  vh_set_field_limits(s->gobox, s->gfolim, incOuter);
  vh_set_field_limits(s->ggbox, s->gfglim, incGhost);
}

static void alloc_grid(vh_solve *s)
{
  real **oneD[12] = {&s->dex, &s->dey, &s->dez, &s->idex, &s->idey, &s->idez,
                     &s->dhx, &s->dhy, &s->dhz, &s->idhx, &s->idhy, &s->idhz};
  for (int a = 0; a < 12; a++) *oneD[a] = vh_alloc(s, s->numCells[a % 3], sizeof(real));
  real **threeD[15] = {&s->Ex, &s->Ey, &s->Ez, &s->Hx, &s->Hy, &s->Hz,
                       &s->alphaEx, &s->alphaEy, &s->alphaEz, &s->betaEx, &s->betaEy, &s->betaEz,
                       &s->gammaHx, &s->gammaHy, &s->gammaHz};
  for (int a = 0; a < 15; a++) *threeD[a] = vh_alloc(s, s->ncell, sizeof(real));
}

/* Primary edges from the mesh lines, extended into PML and ghost cells;
 * secondary edges are the averages of adjacent primaries (grid.c:583-679). */
static void edges_for_axis(vh_solve *s, int a, real *de, real *ide, real *dh, real *idh, const real *L)
{
  int lo = 2 * a, hi = lo + 1;
  int gol = s->gobox[lo], goh = s->gobox[hi], gil = s->gibox[lo], gih = s->gibox[hi];
  for (int i = gol; i <= gil - 1; i++) de[i] = L[1] - L[0];
  for (int i = gil; i <= gih - 1; i++) de[i] = L[i - gil + 1] - L[i - gil];
  for (int i = gih; i <= goh - 1; i++) de[i] = L[gih - gil] - L[gih - gil - 1];
  de[gol - 1] = de[gol];
  de[goh] = de[goh - 1];
  for (int i = gol - 1; i <= goh; i++) {
    ide[i] = 1.0 / de[i];
    if (de[i] < s->dmin[a]) s->dmin[a] = de[i];
    if (de[i] > s->dmax[a]) s->dmax[a] = de[i];
  }
  for (int i = gol; i <= goh; i++) dh[i] = 0.5 * (de[i] + de[i - 1]);
  dh[gol - 1] = dh[gol];
  for (int i = gol; i <= goh; i++) {
    idh[i] = 1.0 / dh[i];
    if (dh[i] < s->dmin[a]) s->dmin[a] = dh[i];
    if (dh[i] > s->dmax[a]) s->dmax[a] = dh[i];
  }
  idh[gol - 1] = idh[gol];
}

static void init_cell_edges(vh_solve *s)
{
  real maxv;
#if defined(VH_REAL_IS_FLOAT)
  maxv = 3.40282347e+38F;
#else
  maxv = 1.7976931348623157e+308;
#endif
  for (int a = 0; a < 3; a++) { s->dmin[a] = maxv; s->dmax[a] = 0.0; }
  edges_for_axis(s, 0, s->dex, s->idex, s->dhx, s->idhx, s->xlines);
  edges_for_axis(s, 1, s->dey, s->idey, s->dhy, s->idhy, s->ylines);
  edges_for_axis(s, 2, s->dez, s->idez, s->dhz, s->idhz, s->zlines);
  s->dt = s->courantNumber / s->c0 / sqrt((1.0 / s->dmin[0]) * (1.0 / s->dmin[0]) +
                                          (1.0 / s->dmin[1]) * (1.0 / s->dmin[1]) +
                                          (1.0 / s->dmin[2]) * (1.0 / s->dmin[2]));
}

#define GRID_TYPE_TOL 1e-5

static void set_grid_type(vh_solve *s)
{
  bool uni[3] = {true, true, true};
  real *de[3] = {s->dex, s->dey, s->dez};
  for (int a = 0; a < 3; a++)
    for (int i = s->gibox[2 * a] + 1; i <= s->gibox[2 * a + 1] - 1; i++)
      if (de[a][i] - de[a][i - 1] > GRID_TYPE_TOL) uni[a] = false;
  if (uni[0] && uni[1] && uni[2]) {
    if ((fabs(s->dex[s->gibox[XLO]] - s->dey[s->gibox[YLO]]) < GRID_TYPE_TOL) &&
        (fabs(s->dey[s->gibox[YLO]] - s->dez[s->gibox[ZLO]]) < GRID_TYPE_TOL))
      s->gridType = GT_CUBIC;
    else
      s->gridType = GT_UNIFORM;
    s->duni[0] = s->dex[s->gibox[XLO]];
    s->duni[1] = s->dey[s->gibox[YLO]];
    s->duni[2] = s->dez[s->gibox[ZLO]];
  } else {
    s->gridType = GT_NONUNIFORM;
    s->duni[0] = s->duni[1] = s->duni[2] = -1.0;
  }
}

static real phase_velocity_func(vh_solve *s, real k, real A[3], real B)
{
  real tmp[3], func, der;
  tmp[0] = sin(A[0] * k) / s->duni[0];
  tmp[1] = sin(A[1] * k) / s->duni[1];
  tmp[2] = sin(A[2] * k) / s->duni[2];
  func = tmp[0] * tmp[0] + tmp[1] * tmp[1] + tmp[2] * tmp[2] - B * B;
  der = A[0] * sin(2 * A[0] * k) / (s->duni[0] * s->duni[0]) +
        A[1] * sin(2 * A[1] * k) / (s->duni[1] * s->duni[1]) +
        A[2] * sin(2 * A[2] * k) / (s->duni[2] * s->duni[2]);
  return func / der;
}

/* Newton iteration for the numerical phase velocity (grid.c:1593-1617). */
real vh_numerical_phase_velocity(vh_solve *s, real theta, real phi)
{
  real A[3], B, w, k;
  w = 2.0 * s->pi * (1.0 / s->dt / 23);
  A[0] = 0.5 * s->duni[0] * sin(theta) * cos(phi);
  A[1] = 0.5 * s->duni[1] * sin(theta) * sin(phi);
  A[2] = 0.5 * s->duni[2] * cos(theta);
  B = (sin(0.5 * w * s->dt) / (s->c0 * s->dt));
  k = w / s->c0;
  for (int i = 0; i < 10; i++) k = k - phase_velocity_func(s, k, A, B);
  return w / k;
}

void vh_field_physical_location(const vh_solve *s, real r[3], int field, int ig, int jg, int kg)
{
  int im = ig - s->gibox[XLO], jm = jg - s->gibox[YLO], km = kg - s->gibox[ZLO];
  r[0] = s->xlines[im];
  r[1] = s->ylines[jm];
  r[2] = s->zlines[km];
  switch (field) {
  case EX: r[0] = s->xlines[im] + 0.5 * s->dex[ig]; break;
  case EY: r[1] = s->ylines[jm] + 0.5 * s->dey[jg]; break;
  case EZ: r[2] = s->zlines[km] + 0.5 * s->dez[kg]; break;
  case HX: r[1] = s->ylines[jm] + 0.5 * s->dey[jg]; r[2] = s->zlines[km] + 0.5 * s->dez[kg]; break;
  case HY: r[0] = s->xlines[im] + 0.5 * s->dex[ig]; r[2] = s->zlines[km] + 0.5 * s->dez[kg]; break;
  case HZ: r[0] = s->xlines[im] + 0.5 * s->dex[ig]; r[1] = s->ylines[jm] + 0.5 * s->dey[jg]; break;
  }
}

void vh_field_index_location(real ijk[3], int field, int ig, int jg, int kg)
{
  ijk[0] = ig;
  ijk[1] = jg;
  ijk[2] = kg;
  switch (field) {
  case EX: ijk[0] = ig + 0.5; break;
  case EY: ijk[1] = jg + 0.5; break;
  case EZ: ijk[2] = kg + 0.5; break;
  case HX: ijk[1] = jg + 0.5; ijk[2] = kg + 0.5; break;
  case HY: ijk[0] = ig + 0.5; ijk[2] = kg + 0.5; break;
  case HZ: ijk[0] = ig + 0.5; ijk[1] = jg + 0.5; break;
  }
}

/* ---------------------------------------------------------------------- */
/* Media (medium.c:381-504, grid.c:1726-1873).                             */
/* ---------------------------------------------------------------------- */

void vh_calc_coeff(vh_solve *s, real *alpha, real *beta, real *gamma, double complex *dalpha,
                   double complex *dbeta, real dt, real eps_r, real sigma, real mu_r, int numPoles,
                   const double complex *residues, const double complex *poles)
{
  double sum = 0.0;
  for (int p = 0; p < numPoles; p++) {
    dalpha[p] = (1 + 0.5 * dt * poles[p]) / (1 - 0.5 * dt * poles[p]);
    dbeta[p] = s->eps0 * residues[p] / (1 - 0.5 * dt * poles[p]);
    sum = sum + creal(dbeta[p]);
  }
  *alpha = (2.0 * eps_r * s->eps0 + 2.0 * sum * dt - dt * sigma) /
           (2.0 * eps_r * s->eps0 + 2.0 * sum * dt + dt * sigma);
  *beta = (2.0 * dt) / (2.0 * eps_r * s->eps0 + 2.0 * sum * dt + dt * sigma);
  *gamma = dt / (mu_r * s->mu0);
}

static void init_media(vh_solve *s)
{
  for (int m = 0; m < s->numMedia; m++) {
    vh_medium *md = &s->media[m];
    if (md->type == MT_DEBYE) {
      md->dalpha = vh_alloc(s, md->numPoles, sizeof(double complex));
      md->dbeta = vh_alloc(s, md->numPoles, sizeof(double complex));
    }
    vh_calc_coeff(s, &md->alpha, &md->beta, &md->gamma, md->dalpha, md->dbeta, s->dt, md->eps_r,
                  md->sigma, md->mu_r, md->numPoles, md->residues, md->poles);
  }
  s->media[MT_FREE_SPACE].alpha = 1.0;
  s->media[MT_FREE_SPACE].beta = s->dt / s->eps0;
  s->media[MT_FREE_SPACE].gamma = s->dt / s->mu0;
  s->media[MT_PEC].alpha = -1.0;
  s->media[MT_PEC].beta = 0.0;
  s->media[MT_PEC].gamma = s->dt / s->mu0;
}

void vh_set_medium_on_grid(vh_solve *s, const int gbbox[6], int medium, FaceMask mask)
{
  bool inc[6];
  int lim[6][6];
  vh_face_mask_to_bool(inc, mask);
  if (vh_bbox_type(gbbox) == BB_POINT) return;
  vh_set_field_limits(gbbox, lim, inc);
  real alpha = s->media[medium].alpha, beta = s->media[medium].beta, gamma = s->media[medium].gamma;
  real *A[3] = {s->alphaEx, s->alphaEy, s->alphaEz}, *B[3] = {s->betaEx, s->betaEy, s->betaEz};
  real *G[3] = {s->gammaHx, s->gammaHy, s->gammaHz};
  for (int f = EX; f <= HZ; f++)
    for (int i = lim[f][XLO]; i <= lim[f][XHI]; i++)
      for (int j = lim[f][YLO]; j <= lim[f][YHI]; j++)
        for (int k = lim[f][ZLO]; k <= lim[f][ZHI]; k++) {
          size_t id = VH_IDX(s, i, j, k);
          if (f < HX) { A[f][id] = alpha; B[f][id] = beta; }
          else G[f - HX][id] = gamma;
        }
}

/* ---------------------------------------------------------------------- */
/* Blocks, Debye blocks, lines (block.c:148-228, debye.c:76-196,           */
/* line.c:165-200).                                                        */
/* ---------------------------------------------------------------------- */

static void init_blocks(vh_solve *s)
{
  int numDebye = 0;
  for (int b = 0; b < s->numBlocks; b++) {
    vh_block *blk = &s->blocks[b];
    int gb[6];
    vh_offset_bbox(gb, blk->mbbox, s->gibox);
    int type = s->media[blk->medium].type;
    if (type == MT_DEBYE) numDebye++;
    vh_set_medium_on_grid(s, gb, blk->medium, blk->mask);
  }
  s->debye = vh_alloc(s, numDebye, sizeof(vh_debye));
  s->numDebye = 0;
  for (int b = 0; b < s->numBlocks; b++) {
    vh_block *blk = &s->blocks[b];
    if (s->media[blk->medium].type != MT_DEBYE) continue;
    vh_debye *d = &s->debye[s->numDebye++];
    int gb[6], lim[6][6];
    bool inc[6];
    vh_offset_bbox(gb, blk->mbbox, s->gibox);
    vh_face_mask_to_bool(inc, blk->mask);
    vh_set_field_limits(gb, lim, inc);
    d->medium = blk->medium;
    int np = s->media[blk->medium].numPoles;
    for (int f = EX; f <= EZ; f++) {
      for (int q = 0; q < 6; q++) d->flim[f][q] = lim[f][q];
      size_t ni = (size_t)(lim[f][XHI] - lim[f][XLO] + 1), nj = (size_t)(lim[f][YHI] - lim[f][YLO] + 1),
             nk = (size_t)(lim[f][ZHI] - lim[f][ZLO] + 1);
      d->n[f] = ni * nj * nk;
      d->J[f] = vh_alloc(s, d->n[f] * (size_t)(np > 0 ? np : 1), sizeof(double complex));
      d->lastE[f] = vh_alloc(s, d->n[f], sizeof(real));
    }
  }
}

static void init_lines(vh_solve *s)
{
  for (int l = 0; l < s->numLinesTW; l++) {
    int gb[6];
    vh_offset_bbox(gb, s->lines[l].mbbox, s->gibox);
    int type = s->wireTypes[s->lines[l].wire];
    vh_set_medium_on_grid(s, gb, type == 0 ? MT_PEC : MT_FREE_SPACE, FACE_MASK_ALL);
  }
}

/* ---------------------------------------------------------------------- */
/* SIBC boundaries and surfaces (sibc.c:126-397, filter.c:142-368).        */
/* ---------------------------------------------------------------------- */

void vh_rc_from_pole_residue(vh_rcfilter *rc, int numPoles, double asymp, const double complex *poles,
                             const double complex *residues, double T)
{
  rc->numPoles = numPoles;
  rc->asymp = asymp;
  size_t n = numPoles > 0 ? (size_t)numPoles : 1;
  rc->q0 = calloc(n, sizeof(double complex));
  rc->q1 = calloc(n, sizeof(double complex));
  rc->q2 = calloc(n, sizeof(double complex));
  rc->poles = calloc(n, sizeof(double complex));
  rc->residues = calloc(n, sizeof(double complex));
  for (int k = 0; k < numPoles; k++) {
    rc->poles[k] = poles[k];
    rc->residues[k] = residues[k];
    double complex alpha = residues[k] / poles[k];
    double complex beta = poles[k] * T;
    rc->q0[k] = cexp(beta);
    rc->q1[k] = alpha / beta * (1.0 + (beta - 1.0) * cexp(beta));
    rc->q2[k] = alpha / beta * (cexp(beta) - beta - 1.0);
  }
}

static void tport_s2z(vh_solve *s, real Z[2][2], real S[2][2])
{
  real DeltaS;
  DeltaS = (1.0 - S[0][0]) * (1.0 - S[1][1]) - S[0][1] * S[1][0];
  Z[0][0] = ((1.0 + S[0][0]) * (1.0 - S[1][1]) + S[0][1] * S[1][0]) / DeltaS * s->eta0;
  Z[0][1] = 2.0 * S[0][1] * s->eta0 / DeltaS;
  Z[1][0] = 2.0 * S[1][0] * s->eta0 / DeltaS;
  Z[1][1] = ((1.0 - S[0][0]) * (1.0 + S[1][1]) + S[0][1] * S[1][0]) / DeltaS * s->eta0;
}

static bool is_passive_s(real S[2][2])
{
  bool e00 = (1 - (S[0][0] * S[0][0] + S[1][0] * S[1][0])) >= 0.0;
  bool e01 = (S[0][0] * S[0][1] - S[1][0] * S[1][1]) >= 0.0;
  bool e10 = (S[0][1] * S[0][0] - S[1][1] * S[1][0]) >= 0.0;
  bool e11 = (1 - (S[0][1] * S[0][1] + S[1][1] * S[1][1])) >= 0.0;
  return e00 && e01 && e10 && e11;
}

static void init_sibc_boundary(vh_solve *s, vh_boundary *b)
{
  double T = (double)s->dt;
  if (strlen(b->fileName) == 0) {
    real Z_TM[2][2], Z_TE[2][2];
    if (!is_passive_s(b->S_TM) || !is_passive_s(b->S_TE))
      vh_fail(s, "passivity violation for SIBC boundary %s", b->name);
    tport_s2z(s, Z_TM, b->S_TM);
    tport_s2z(s, Z_TE, b->S_TE);
    for (int p = 0; p < 4; p++)
      for (int q = 0; q < 4; q++) {
        double asymp = 0.0;
        if (p < 2 && q < 2) asymp = Z_TM[p][q];
        if (p >= 2 && q >= 2) asymp = Z_TE[p - 2][q - 2];
        vh_rc_from_pole_residue(&b->rc[4 * p + q], 0, asymp, NULL, NULL, T);
      }
  } else {
    FILE *fp = vh_open_input(s, b->fileName, "r");
    int m, n;
    if (!fp) vh_fail(s, "cannot open pole-residue matrix file %s", b->fileName);
    if (fscanf(fp, "%d %d", &m, &n) != 2) vh_fail(s, "error reading extents from %s", b->fileName);
    if (!((m == 4 && n == 4) || (m == 2 && n == 2)))
      vh_fail(s, "pole-residue model in %s must be 2x2 or 4x4", b->fileName);
    int np[4][4] = {{0}};
    double as[4][4] = {{0}};
    double complex *po[4][4] = {{0}}, *re[4][4] = {{0}};
    for (int r = 0; r < m; r++)
      for (int c = 0; c < n; c++) {
        int nump = 0;
        double asymp = 0.0;
        if (fscanf(fp, "%d", &nump) != 1 || fscanf(fp, "%lf", &asymp) != 1)
          vh_fail(s, "error reading filter (%d,%d) from %s", r, c, b->fileName);
        np[r][c] = nump;
        as[r][c] = asymp;
        po[r][c] = vh_alloc(s, nump, sizeof(double complex));
        re[r][c] = vh_alloc(s, nump, sizeof(double complex));
        for (int k = 0; k < nump; k++) {
          double pr, pi_, rr, ri;
          if (fscanf(fp, "%lf %lf %lf %lf", &pr, &pi_, &rr, &ri) != 4)
            vh_fail(s, "error reading poles of filter (%d,%d) from %s", r, c, b->fileName);
          po[r][c][k] = pr + I * pi_;
          re[r][c][k] = rr + I * ri;
        }
      }
    fclose(fp);
    for (int p = 0; p < 4; p++)
      for (int q = 0; q < 4; q++) {
        int sp = p, sq = q, use = 1;
        if (m == 2) {
          /* isotropic: TM block and TE block both from the 2x2 model */
          if ((p < 2) != (q < 2)) use = 0;
          sp = p % 2;
          sq = q % 2;
        }
        if (use)
          vh_rc_from_pole_residue(&b->rc[4 * p + q], np[sp][sq], as[sp][sq], po[sp][sq], re[sp][sq], T);
        else
          vh_rc_from_pole_residue(&b->rc[4 * p + q], 0, 0.0, NULL, NULL, T);
      }
    for (int r = 0; r < m; r++)
      for (int c = 0; c < n; c++) { free(po[r][c]); free(re[r][c]); }
  }
  b->hasRc = 1;
}

static void init_boundaries(vh_solve *s)
{
  for (int b = 0; b < s->numBounds; b++)
    if (s->bounds[b].type == BT_SIBC) init_sibc_boundary(s, &s->bounds[b]);
}

#define SFACE(s, i, j, k, d) (((VH_IDX(s, i, j, k)) * 3) + (size_t)(d))

static void init_internal_surfaces(vh_solve *s)
{
  int numSibc = 0;
  for (int n = 0; n < s->numSurfaces; n++) {
    vh_surface *sf = &s->surfaces[n];
    int type = s->bounds[sf->boundary].type;
    switch (type) {
    case BT_PEC:
      vh_offset_bbox(sf->gbbox, sf->mbbox, s->gibox);
      vh_set_medium_on_grid(s, sf->gbbox, MT_PEC, FACE_MASK_ALL);
      break;
    case BT_FREE_SPACE:
      vh_offset_bbox(sf->gbbox, sf->mbbox, s->gibox);
      vh_set_medium_on_grid(s, sf->gbbox, MT_FREE_SPACE, FACE_MASK_ALL);
      break;
    case BT_SIBC:
      vh_offset_bbox(sf->gbbox, sf->mbbox, s->gibox);
      vh_set_medium_on_grid(s, sf->gbbox, MT_PEC, FACE_MASK_ALL);
      numSibc++;
      break;
    default: break;
    }
  }
  if (numSibc == 0) return;
  s->sibc = vh_alloc(s, numSibc, sizeof(vh_sibc));
  bool *isFace = vh_alloc(s, s->ncell * 3, sizeof(bool));
  for (int n = 0; n < s->numSurfaces; n++) {
    vh_surface *sf = &s->surfaces[n];
    if (s->bounds[sf->boundary].type != BT_SIBC) continue;
    vh_sibc *sb = &s->sibc[s->numSibc++];
    sb->boundary = sf->boundary;
    sb->orientation = sf->orient;
    sb->normal = vh_bbox_direction(sf->mbbox);
    sb->cosa = cos(s->pi * sf->angle / 180.0);
    sb->sina = sin(s->pi * sf->angle / 180.0);
    vh_offset_bbox(sb->gbbox, sf->mbbox, s->gibox);
    sb->gbbox[2 * sb->normal + 1] += 1;
    size_t ni = (size_t)(sb->gbbox[XHI] - sb->gbbox[XLO]), nj = (size_t)(sb->gbbox[YHI] - sb->gbbox[YLO]),
           nk = (size_t)(sb->gbbox[ZHI] - sb->gbbox[ZLO]);
    sb->nfaces = ni * nj * nk;
    int maxP = 1;
    for (int f = 0; f < 16; f++)
      if (s->bounds[sb->boundary].rc[f].numPoles > maxP) maxP = s->bounds[sb->boundary].rc[f].numPoles;
    sb->maxPoles = maxP;
    sb->zeta = vh_alloc(s, sb->nfaces * 16 * (size_t)maxP, sizeof(double complex));
    sb->old = vh_alloc(s, sb->nfaces * 16, sizeof(double));
    sb->Etan = vh_alloc(s, sb->nfaces * 4, sizeof(real));
    sb->isAdjA = vh_alloc(s, sb->nfaces * 4, sizeof(bool));
    sb->isAdjB = vh_alloc(s, sb->nfaces * 4, sizeof(bool));
    for (int i = sb->gbbox[XLO]; i < sb->gbbox[XHI]; i++)
      for (int j = sb->gbbox[YLO]; j < sb->gbbox[YHI]; j++)
        for (int k = sb->gbbox[ZLO]; k < sb->gbbox[ZHI]; k++) isFace[SFACE(s, i, j, k, sb->normal)] = true;
  }
  for (int n = 0; n < s->numSibc; n++) {
    vh_sibc *sb = &s->sibc[n];
    size_t f = 0;
    for (int i = sb->gbbox[XLO]; i < sb->gbbox[XHI]; i++)
      for (int j = sb->gbbox[YLO]; j < sb->gbbox[YHI]; j++)
        for (int k = sb->gbbox[ZLO]; k < sb->gbbox[ZHI]; k++, f++) {
          bool *A = &sb->isAdjA[4 * f], *B = &sb->isAdjB[4 * f];
          switch (sb->normal) {
          case XDIR:
            A[0] = isFace[SFACE(s, i - 1, j, k, YDIR)];
            A[1] = isFace[SFACE(s, i - 1, j + 1, k, YDIR)];
            A[2] = isFace[SFACE(s, i - 1, j, k, ZDIR)];
            A[3] = isFace[SFACE(s, i - 1, j, k + 1, ZDIR)];
            B[0] = isFace[SFACE(s, i, j, k, YDIR)];
            B[1] = isFace[SFACE(s, i, j + 1, k, YDIR)];
            B[2] = isFace[SFACE(s, i, j, k, ZDIR)];
            B[3] = isFace[SFACE(s, i, j + 1, k + 1, ZDIR)];
            break;
          case YDIR:
            A[0] = isFace[SFACE(s, i, j - 1, k, ZDIR)];
            A[1] = isFace[SFACE(s, i, j - 1, k + 1, ZDIR)];
            A[2] = isFace[SFACE(s, i, j - 1, k, XDIR)];
            A[3] = isFace[SFACE(s, i + 1, j - 1, k, XDIR)];
            B[0] = isFace[SFACE(s, i, j, k, ZDIR)];
            B[1] = isFace[SFACE(s, i, j, k + 1, ZDIR)];
            B[2] = isFace[SFACE(s, i, j, k, XDIR)];
            B[3] = isFace[SFACE(s, i + 1, j, k + 1, XDIR)];
            break;
          case ZDIR:
            A[0] = isFace[SFACE(s, i, j, k - 1, XDIR)];
            A[1] = isFace[SFACE(s, i + 1, j, k - 1, XDIR)];
            A[2] = isFace[SFACE(s, i, j, k - 1, YDIR)];
            A[3] = isFace[SFACE(s, i, j + 1, k - 1, YDIR)];
            B[0] = isFace[SFACE(s, i, j, k, XDIR)];
            B[1] = isFace[SFACE(s, i + 1, j, k, XDIR)];
            B[2] = isFace[SFACE(s, i, j, k, YDIR)];
            B[3] = isFace[SFACE(s, i + 1, j + 1, k, YDIR)];
            break;
          }
        }
  }
  free(isFace);
}

/* ---------------------------------------------------------------------- */
/* External surfaces: PEC backing, UPML, Mur (surface.c:409-458,           */
/* pml.c:90-629, mur.c:101-131).                                           */
/* ---------------------------------------------------------------------- */

static void init_external_pec(vh_solve *s)
{
  /* Boundary type -> medium (boundary.c:43): only PML, PEC and SIBC set PEC. */
  for (int f = 0; f < 6; f++) {
    int t = vh_outer_type(s, f);
    if (t == BT_PML || t == BT_PEC || t == BT_SIBC) {
      int bb[6];
      vh_face_of_bbox(bb, s->gobox, f);
      vh_set_medium_on_grid(s, bb, MT_PEC, FACE_MASK_ALL);
    }
  }
}

static real pml_sigma_profile(vh_solve *s, real x, real totalDepth, real meshSize, int order, real n_eff,
                              real refCoeff)
{
  real sigmaMax;
  if (refCoeff > 0)
    sigmaMax = -(order + 1.0) / (2.0 * s->eta0 * n_eff * totalDepth) * log(refCoeff);
  else
    sigmaMax = 4.0 * (order + 1.0) / (5.0 * s->eta0 * n_eff * meshSize);
  return sigmaMax * pow(x, order);
}

static real pml_kappa_profile(real x, int order, real kmax)
{
  return 1.0 + (kmax - 1.0) * pow(x, order);
}

static void set_profile(vh_solve *s, const real *de, const real *dh, int low, int high, int dir, int order,
                        real n_eff, real refCoeff, real kmax, real *gd, real *kd, real *bd, real *ad,
                        real *gh, real *kh, real *bh, real *ah, real *ibd, real *ibh)
{
  real total = 0.0, depthE, depthH, sprof, kprof;
  real dt = s->dt;
  for (int i = low; i <= high; i++) total = total + de[i];
  if (dir == -1) {
    depthE = total;
    depthH = total - 0.5 * de[low];
  } else {
    depthE = 0.0;
    depthH = 0.5 * de[low];
  }
  for (int i = low; i <= high; i++) {
    sprof = 0.5 * dt / s->eps0 * pml_sigma_profile(s, depthE / total, total, de[low], order, n_eff, refCoeff);
    kprof = pml_kappa_profile(depthE / total, order, kmax);
    gh[i] = sprof;
    kh[i] = kprof;
    bd[i] = 1.0 / (kprof + sprof);
    ibd[i] = kprof + sprof;
    ad[i] = (kprof - sprof) / (kprof + sprof);
    depthE = depthE + dir * de[i];
  }
  if (dir == +1) {
    sprof = 0.5 * dt / s->eps0 * pml_sigma_profile(s, depthE / total, total, de[low], order, n_eff, refCoeff);
    kprof = pml_kappa_profile(depthE / total, order, kmax);
    gh[high + 1] = sprof;
    kh[high + 1] = kprof;
    bd[high + 1] = 1.0 / (kprof + sprof);
    ibd[high + 1] = kprof + sprof;
    ad[high + 1] = (kprof - sprof) / (kprof + sprof);
  }
  for (int i = low; i <= high; i++) {
    sprof = 0.5 * dt / s->eps0 * pml_sigma_profile(s, depthH / total, total, de[low], order, n_eff, refCoeff);
    kprof = pml_kappa_profile(depthH / total, order, kmax);
    gd[i] = sprof;
    kd[i] = kprof;
    bh[i] = 1.0 / (kprof + sprof);
    ibh[i] = kprof + sprof;
    ah[i] = (kprof - sprof) / (kprof + sprof);
    depthH = depthH + dir * dh[i];
  }
}

static void init_pml(vh_solve *s)
{
  /* Region limits (pml.c:133-195). */
  static const bool incB[6][6] = {{1, 0, 1, 1, 1, 1}, {0, 1, 1, 1, 1, 1}, {1, 1, 1, 0, 1, 1},
                                  {1, 1, 0, 1, 1, 1}, {1, 1, 1, 1, 1, 0}, {1, 1, 1, 1, 0, 1}};
  int *go = s->gobox, *gi = s->gibox;
  for (int r = 0; r < 6; r++) {
    int a = r >> 1;
    for (int f = 0; f < 6; f++) {
      int fa = f >> 1;
      if (fa == a) {
        /* along the region's own axis: between outer and inner box */
        s->pbox[r][f] = (r & 1) ? ((f & 1) ? go[f] : gi[f + 1]) : ((f & 1) ? gi[f - 1] : go[f]);
      } else if (fa < a) {
        /* axes handled by an earlier (lower-axis) region pair: inner extent */
        s->pbox[r][f] = gi[f];
      } else {
        s->pbox[r][f] = go[f];
      }
    }
  }
  /* X slabs span the outer y/z extent, Y slabs the inner x and outer z,
   * Z slabs the inner x and y. */
  for (int r = 0; r < 6; r++)
    vh_set_field_limits(s->pbox[r], s->fplim[r], incB[r]);

  size_t nx = (size_t)s->numCells[0], ny = (size_t)s->numCells[1], nz = (size_t)s->numCells[2];
  for (int r = 0; r < 6; r++) {
    if (vh_outer_type(s, r) != BT_PML) continue;
    real **arrs[9] = {&s->Px[r], &s->Py[r], &s->Pz[r], &s->PPx[r], &s->PPy[r], &s->PPz[r],
                      &s->Bx[r], &s->By[r], &s->Bz[r]};
    static const int comp[9] = {EX, EY, EZ, EX, EY, EZ, HX, HY, HZ};
    for (int q = 0; q < 9; q++) {
      int *l = s->fplim[r][comp[q]];
      long n = (long)(l[XHI] - l[XLO] + 1) * (l[YHI] - l[YLO] + 1) * (l[ZHI] - l[ZLO] + 1);
      *arrs[q] = vh_alloc(s, n > 0 ? (size_t)n : 1, sizeof(real));
    }
  }
  real **prof[30] = {&s->adx, &s->bdx, &s->gdx, &s->kdx, &s->ahx, &s->bhx, &s->ghx, &s->khx, &s->ibdx, &s->ibhx,
                     &s->ady, &s->bdy, &s->gdy, &s->kdy, &s->ahy, &s->bhy, &s->ghy, &s->khy, &s->ibdy, &s->ibhy,
                     &s->adz, &s->bdz, &s->gdz, &s->kdz, &s->ahz, &s->bhz, &s->ghz, &s->khz, &s->ibdz, &s->ibhz};
  size_t nn[3] = {nx, ny, nz};
  for (int q = 0; q < 30; q++) *prof[q] = vh_alloc(s, nn[q / 10], sizeof(real));
  for (int a = 0; a < 3; a++) {
    real *kd = *prof[10 * a + 3], *kh = *prof[10 * a + 7], *gd = *prof[10 * a + 2], *gh = *prof[10 * a + 6];
    real *bd = *prof[10 * a + 1], *bh = *prof[10 * a + 5], *ibd = *prof[10 * a + 8], *ibh = *prof[10 * a + 9];
    real *ad = *prof[10 * a + 0], *ah = *prof[10 * a + 4];
    for (int i = go[2 * a] - 1; i <= go[2 * a + 1]; i++) {
      kd[i] = 1.0; kh[i] = 1.0; gd[i] = 0.0; gh[i] = 0.0; bd[i] = 1.0;
      bh[i] = 1.0; ibd[i] = 1.0; ibh[i] = 1.0; ad[i] = 1.0; ah[i] = 1.0;
    }
  }
  const real *de[3] = {s->dex, s->dey, s->dez}, *dh[3] = {s->dhx, s->dhy, s->dhz};
  for (int f = 0; f < 6; f++) {
    if (vh_outer_type(s, f) != BT_PML) continue;
    vh_boundary *b = &s->bounds[s->extBoundary[f]];
    int a = f >> 1;
    int low = (f & 1) ? gi[f] : go[f];
    int high = (f & 1) ? go[f] - 1 : gi[f] - 1;
    int dir = (f & 1) ? +1 : -1;
    set_profile(s, de[a], dh[a], low, high, dir, b->order, b->n_eff, b->refCoeff, b->kmax,
                *prof[10 * a + 2], *prof[10 * a + 3], *prof[10 * a + 1], *prof[10 * a + 0],
                *prof[10 * a + 6], *prof[10 * a + 7], *prof[10 * a + 5], *prof[10 * a + 4],
                *prof[10 * a + 8], *prof[10 * a + 9]);
  }

  /* Carry boundary materials into the slabs, Z then Y then X (pml.c:457-629). */
  static const bool incM[6][6] = {{0, 0, 1, 1, 1, 1}, {0, 0, 1, 1, 1, 1}, {1, 1, 0, 0, 1, 1},
                                  {1, 1, 0, 0, 1, 1}, {1, 1, 1, 1, 0, 0}, {1, 1, 1, 1, 0, 0}};
  int fpm[6][6][6];
  for (int r = 0; r < 6; r++) vh_set_field_limits(s->pbox[r], fpm[r], incM[r]);
  real *A[3] = {s->alphaEx, s->alphaEy, s->alphaEz}, *B[3] = {s->betaEx, s->betaEy, s->betaEz};
  real *G[3] = {s->gammaHx, s->gammaHy, s->gammaHz};
  static const int order3[3] = {2, 1, 0};
  for (int oa = 0; oa < 3; oa++) {
    int a = order3[oa];
    for (int r = 2 * a; r <= 2 * a + 1; r++) {
      int offset = (r & 1) ? -1 : +1;
      for (int field = EX; field <= HZ; field++) {
        /* Tangential E and normal H are copied from the interface plane;
         * normal E and tangential H from the plane one cell inside. */
        int src = gi[r] + (vh_field_is_in_boundary(field, r) ? 0 : offset);
        int *l = fpm[r][field];
        for (int i = l[XLO]; i <= l[XHI]; i++)
          for (int j = l[YLO]; j <= l[YHI]; j++)
            for (int k = l[ZLO]; k <= l[ZHI]; k++) {
              int si = i, sj = j, sk = k;
              if (a == 0) si = src; else if (a == 1) sj = src; else sk = src;
              size_t d = VH_IDX(s, i, j, k), o = VH_IDX(s, si, sj, sk);
              if (field < HX) { A[field][d] = A[field][o]; B[field][d] = B[field][o]; }
              else G[field - HX][d] = G[field - HX][o];
            }
      }
    }
  }
}

static void init_mur(vh_solve *s)
{
  for (int f = 0; f < 6; f++) {
    if (vh_outer_type(s, f) != BT_MUR) continue;
    bool inc[6] = {true, true, true, true, true, true};
    vh_face_of_bbox(s->murbox[f], s->gibox, f);
    for (int g = 0; g < 6; g++)
      if ((g >> 1) != (f >> 1) && vh_outer_type(s, g) == BT_MUR) inc[g] = false;
    vh_set_field_limits(s->murbox[f], s->fmlim[f], inc);
  }
  real dt = s->dt;
  /* The high faces index the edge array with box entry [face-1], i.e. the
   * low entry of the face box, which equals the high one (mur.c:122-127). */
  s->zeta[XLO] = (s->c0 * dt - s->dex[s->murbox[XLO][XLO]]) / (s->c0 * dt + s->dex[s->murbox[XLO][XLO]]);
  s->zeta[XHI] = (s->c0 * dt - s->dex[s->murbox[XHI][XLO]]) / (s->c0 * dt + s->dex[s->murbox[XHI][XLO]]);
  s->zeta[YLO] = (s->c0 * dt - s->dey[s->murbox[YLO][YLO]]) / (s->c0 * dt + s->dey[s->murbox[YLO][YLO]]);
  s->zeta[YHI] = (s->c0 * dt - s->dey[s->murbox[YHI][YLO]]) / (s->c0 * dt + s->dey[s->murbox[YHI][YLO]]);
  s->zeta[ZLO] = (s->c0 * dt - s->dez[s->murbox[ZLO][ZLO]]) / (s->c0 * dt + s->dez[s->murbox[ZLO][ZLO]]);
  s->zeta[ZHI] = (s->c0 * dt - s->dez[s->murbox[ZHI][ZLO]]) / (s->c0 * dt + s->dez[s->murbox[ZHI][ZLO]]);
}

/* ---------------------------------------------------------------------- */
/* Waveforms (waveform.c:256-347, 511-805).                                */
/* ---------------------------------------------------------------------- */

static void create_splines(vh_solve *s, vh_waveform *w, real deriv1, real deriv2)
{
  unsigned long n = w->tableSize;
  real *T = w->table;
#define TB(i, c) T[3 * (i) + (c)]
  real p, qnm1, sign, unm1;
  if (n < 2UL) vh_fail(s, "insufficient points for spline evaluation in %s", w->fileName);
  for (unsigned long i = 1UL; i < n; i++)
    if (TB(i, 0) - TB(i - 1, 0) <= 0.0) vh_fail(s, "time data in %s is not increasing", w->fileName);
  real *u = vh_alloc(s, n - 1UL, sizeof(real));
  if (deriv1 > 0.99e30) {
    TB(0, 2) = 0.0;
    u[0] = 0.0;
  } else {
    TB(0, 2) = -0.5;
    u[0] = (3.0 / (TB(1, 0) - TB(0, 0))) * ((TB(1, 1) - TB(0, 1)) / (TB(1, 0) - TB(0, 0)) - deriv1);
  }
  for (unsigned long i = 1UL; i <= n - 2UL; i++) {
    sign = (TB(i, 0) - TB(i - 1, 0)) / (TB(i + 1, 0) - TB(i - 1, 0));
    p = sign * TB(i - 1, 2) + 2.0;
    TB(i, 2) = (sign - 1.0) / p;
    u[i] = (TB(i + 1, 1) - TB(i, 1)) / (TB(i + 1, 0) - TB(i, 0)) - (TB(i, 1) - TB(i - 1, 1)) / (TB(i, 0) - TB(i - 1, 0));
    u[i] = (6.0 * u[i] / (TB(i + 1, 0) - TB(i - 1, 0)) - sign * u[i - 1]) / p;
  }
  if (deriv2 > 0.99e30) {
    qnm1 = 0.0;
    unm1 = 0.0;
  } else {
    qnm1 = 0.5;
    unm1 = (3.0 / (TB(n - 1, 0) - TB(n - 2, 0))) * (deriv2 - (TB(n - 1, 1) - TB(n - 2, 1)) / (TB(n - 1, 0) - TB(n - 2, 0)));
  }
  TB(n - 1, 2) = (unm1 - qnm1 * u[n - 2]) / (qnm1 * TB(n - 2, 2) + 1.0);
  for (unsigned long k = n - 1UL; k-- > 0;) TB(k, 2) = TB(k, 2) * TB(k + 1, 2) + u[k];
  free(u);
#undef TB
}

static void load_external_waveform(vh_solve *s, vh_waveform *w)
{
  FILE *fp = vh_open_input(s, w->fileName, "r");
  double t1, t2;
  unsigned long n = 0;
  if (!fp) vh_fail(s, "cannot open external waveform file %s", w->fileName);
  while (fscanf(fp, "%lf %lf", &t1, &t2) == 2) n++;
  if (!feof(fp)) vh_fail(s, "error reading from %s", w->fileName);
  rewind(fp);
  w->table = vh_alloc(s, 3 * (n ? n : 1), sizeof(real));
  for (unsigned long i = 0; i < n; i++) {
    if (fscanf(fp, "%lf %lf", &t1, &t2) != 2) vh_fail(s, "error reading from %s", w->fileName);
    w->table[3 * i] = (real)t1;
    w->table[3 * i + 1] = (real)t2;
    w->table[3 * i + 2] = 0.0;
  }
  fclose(fp);
  w->tableSize = n;
}

static void init_waveforms(vh_solve *s)
{
  real del_t = s->dt;
  for (int n = 0; n < s->numWaveforms; n++) {
    vh_waveform *w = &s->waveforms[n];
    if (w->size < 0) w->size = 1.0;
    switch (w->type) {
    case WT_GAUSSIAN_PULSE: case WT_DIFFERENTIATED_GAUSSIAN_PULSE: case WT_RICKER_WAVELET:
      if (w->width < 0) w->width = 5.0 * sqrt(2.0) * del_t;
      if (w->delay < 0) w->delay = 40.0 * del_t;
      w->frequency = 0.0;
      break;
    case WT_NARROW_GAUSSIAN_PULSE:
      if (w->width < 0) w->width = 8.0 * del_t;
      if (w->delay < 0) w->delay = 12.0 * del_t;
      w->frequency = 0.0;
      break;
    case WT_MODULATED_GAUSSIAN_PULSE:
      if (w->width < 0) w->width = 20.0 * sqrt(2.0) * del_t;
      if (w->delay < 0) w->delay = 120.0 * del_t;
      if (w->frequency < 0) w->frequency = 0.05 / del_t;
      break;
    case WT_COMPACT_PULSE: case WT_DIFFERENTIATED_COMPACT_PULSE:
      if (w->width < 0) w->width = 20.0 * del_t;
      if (w->delay < 0) w->delay = 0.0;
      w->frequency = 0.0;
      break;
    case WT_MODULATED_COMPACT_PULSE:
      if (w->width < 0) w->width = 80.0 * del_t;
      if (w->delay < 0) w->delay = 0.0;
      if (w->frequency < 0) w->frequency = 0.05 / del_t;
      break;
    case WT_RAMPED_SINUSOID:
      if (w->width < 0) w->width = 20.0 * del_t;
      if (w->delay < 0) w->delay = 0.0;
      if (w->frequency < 0) w->frequency = 0.05 / del_t;
      break;
    case WT_EXTERNAL:
      if (w->delay < 0) w->delay = 0.0;
      load_external_waveform(s, w);
      w->lastIdx = 0UL;
      create_splines(s, w, 0.0, 0.0);
      break;
    }
  }
}

/* ---------------------------------------------------------------------- */
/* Sources (source.c:469-599, 934-1009).                                   */
/* ---------------------------------------------------------------------- */

static void electric_source_size(vh_solve *s, const int gb[6], int dir, real *length, real *area, real side[3])
{
  *length = 0.0; *area = 0.0; side[0] = side[1] = side[2] = 0.0;
  switch (dir) {
  case XDIR:
    for (int i = gb[XLO]; i <= gb[XHI] - 1; i++) *length += s->dex[i];
    for (int j = gb[YLO]; j <= gb[YHI]; j++) side[1] += s->dhy[j];
    for (int k = gb[ZLO]; k <= gb[ZHI]; k++) side[2] += s->dhz[k];
    *area = side[1] * side[2];
    break;
  case YDIR:
    for (int i = gb[XLO]; i <= gb[XHI]; i++) side[0] += s->dhx[i];
    for (int j = gb[YLO]; j <= gb[YHI] - 1; j++) *length += s->dey[j];
    for (int k = gb[ZLO]; k <= gb[ZHI]; k++) side[2] += s->dhz[k];
    *area = side[2] * side[0];
    break;
  case ZDIR:
    for (int i = gb[XLO]; i <= gb[XHI]; i++) side[0] += s->dhx[i];
    for (int j = gb[YLO]; j <= gb[YHI]; j++) side[1] += s->dhy[j];
    for (int k = gb[ZLO]; k <= gb[ZHI] - 1; k++) *length += s->dez[k];
    *area = side[0] * side[1];
    break;
  }
}

static void magnetic_source_size(vh_solve *s, const int gb[6], int dir, real *length, real *area, real side[3])
{
  *length = 0.0; *area = 0.0; side[0] = side[1] = side[2] = 0.0;
  switch (dir) {
  case XDIR:
    for (int i = gb[XLO]; i <= gb[XHI]; i++) *length += s->dhx[i];
    for (int j = gb[YLO]; j <= gb[YHI] - 1; j++) side[1] += s->dey[j];
    for (int k = gb[ZLO]; k <= gb[ZHI] - 1; k++) side[2] += s->dez[k];
    *area = side[1] * side[2];
    break;
  case YDIR:
    for (int i = gb[XLO]; i <= gb[XHI] - 1; i++) side[0] += s->dex[i];
    for (int j = gb[YLO]; j <= gb[YHI]; j++) *length += s->dhy[j];
    for (int k = gb[ZLO]; k <= gb[ZHI] - 1; k++) side[2] += s->dez[k];
    *area = side[2] * side[0];
    break;
  case ZDIR:
    for (int i = gb[XLO]; i <= gb[XHI] - 1; i++) side[0] += s->dex[i];
    for (int j = gb[YLO]; j <= gb[YHI] - 1; j++) side[1] += s->dey[j];
    for (int k = gb[ZLO]; k <= gb[ZHI]; k++) *length += s->dhz[k];
    *area = side[0] * side[1];
    break;
  }
}

#define MIN_RESISTANCE 1e-2

static void init_sources(vh_solve *s)
{
  bool inc[6] = {true, true, true, true, true, true};
  for (int n = 0; n < s->numSources; n++) {
    vh_source *it = &s->sources[n];
    real length, area, side[3], sigma;
    vh_offset_bbox(it->gbbox, it->mbbox, s->gibox);
    vh_set_field_limits(it->gbbox, it->flim, inc);
    switch (it->type) {
    case ST_ELEC_SURF_CURR_DENSITY:
      electric_source_size(s, it->gbbox, it->pol, &length, &area, side);
      it->size = it->size / side[it->normal];
      it->type = ST_ELEC_CURR_DENSITY;
      break;
    case ST_MAGN_SURF_CURR_DENSITY:
      magnetic_source_size(s, it->gbbox, it->pol, &length, &area, side);
      it->size = it->size / side[it->normal];
      it->type = ST_MAGN_CURR_DENSITY;
      break;
    case ST_ELEC_CURRENT:
      electric_source_size(s, it->gbbox, it->pol, &length, &area, side);
      it->size = it->size / area;
      it->type = ST_ELEC_CURR_DENSITY;
      break;
    case ST_MAGN_CURRENT:
      magnetic_source_size(s, it->gbbox, it->pol, &length, &area, side);
      it->size = it->size / area;
      it->type = ST_MAGN_CURR_DENSITY;
      break;
    case ST_ELEC_CURRENT_MOMENT:
      electric_source_size(s, it->gbbox, it->pol, &length, &area, side);
      it->size = it->size / area / length;
      it->type = ST_ELEC_CURR_DENSITY;
      break;
    case ST_MAGN_CURRENT_MOMENT:
      magnetic_source_size(s, it->gbbox, it->pol, &length, &area, side);
      it->size = it->size / area / length;
      it->type = ST_MAGN_CURR_DENSITY;
      break;
    case ST_VOLTAGE:
      electric_source_size(s, it->gbbox, it->pol, &length, &area, side);
      it->size = -it->size / length;
      it->type = ST_EFIELD;
      break;
    case ST_THEVENIN_VOLTAGE:
    case ST_NORTON_CURRENT: {
      if (it->type == ST_NORTON_CURRENT) {
        it->size = it->resistance * it->size;
        it->type = ST_THEVENIN_VOLTAGE;
      }
      char rname[VH_TAG];
      snprintf(rname, sizeof(rname), "__VR_RS_%d__", n + 1);
      int ri = vh_find_medium(s, rname);
      if (ri < 0) vh_fail(s, "internal: lumped resistance medium missing for source %s", it->name);
      electric_source_size(s, it->gbbox, it->pol, &length, &area, side);
      if (it->resistance > MIN_RESISTANCE) {
        it->size = it->size / area / it->resistance;
        sigma = length / (area * it->resistance);
        vh_medium *m = &s->media[ri];
        m->eps_r = 1.0; m->sigma = sigma; m->mu_r = 1.0;
        vh_calc_coeff(s, &m->alpha, &m->beta, &m->gamma, NULL, NULL, s->dt, 1.0, sigma, 1.0, 0, NULL, NULL);
        vh_set_medium_on_grid(s, it->gbbox, ri, FACE_MASK_ALL);
        it->type = ST_ELEC_CURR_DENSITY;
      } else {
        it->size = -0.5 * it->size / length;
        vh_set_medium_on_grid(s, it->gbbox, MT_PEC, FACE_MASK_ALL);
        it->type = ST_EFIELD;
      }
      break;
    }
    default: break;
    }
  }
}

/* ---------------------------------------------------------------------- */
/* Plane waves (planewave.c:277-554, 1226-1296).                           */
/* ---------------------------------------------------------------------- */

static real deg2rad(vh_solve *s, real angle) { return s->pi / 180.0 * angle; }

static void calc_incident_vectors(vh_solve *s, vh_planewave *pw)
{
  real th = deg2rad(s, pw->theta), ph = deg2rad(s, pw->phi), et = deg2rad(s, pw->eta);
  real size = pw->size;
  pw->kinc[0] = sin(th) * cos(ph);
  pw->kinc[1] = sin(th) * sin(ph);
  pw->kinc[2] = cos(th);
  pw->Finc[EX] = size * (cos(et) * sin(ph) - sin(et) * cos(th) * cos(ph));
  pw->Finc[EY] = size * (-cos(et) * cos(ph) - sin(et) * cos(th) * sin(ph));
  pw->Finc[EZ] = size * (sin(et) * sin(th));
  pw->Finc[HX] = size / s->eta0 * (sin(et) * sin(ph) + cos(et) * cos(th) * cos(ph));
  pw->Finc[HY] = size / s->eta0 * (-sin(et) * cos(ph) + cos(et) * cos(th) * sin(ph));
  pw->Finc[HZ] = size / s->eta0 * (-cos(et) * sin(th));
  /* Origin corner from the octant of the incidence direction (eqns 5.61-5.62). */
  int *g = pw->gbbox;
  real theta = pw->theta, phi = pw->phi;
  int zc;
  if (theta >= 0 && theta <= 90) zc = g[ZLO];
  else if (theta > 90 && theta <= 180) zc = g[ZHI];
  else vh_fail(s, "plane wave theta out of range");
  if (phi >= 0 && phi <= 90) { pw->ijk0[0] = g[XLO]; pw->ijk0[1] = g[YLO]; }
  else if (phi > 90 && phi <= 180) { pw->ijk0[0] = g[XHI]; pw->ijk0[1] = g[YLO]; }
  else if (phi > 180 && phi <= 270) { pw->ijk0[0] = g[XHI]; pw->ijk0[1] = g[YHI]; }
  else if (phi > 270 && phi < 360) { pw->ijk0[0] = g[XLO]; pw->ijk0[1] = g[YHI]; }
  else vh_fail(s, "plane wave phi out of range");
  pw->ijk0[2] = zc;
}

static void init_aux_grid(vh_solve *s, vh_planewave *pw)
{
  static const int npml = 10;
  real dt = s->dt, d[3] = {s->duni[0], s->duni[1], s->duni[2]};
  real rel = vh_numerical_phase_velocity(s, 0.0, 0.0) /
             vh_numerical_phase_velocity(s, s->pi * pw->theta / 180.0, s->pi * pw->phi / 180.0);
  pw->betaEyi = dt / (s->eps0 * d[0]) / rel;
  pw->gammaHzi = dt / (s->mu0 * d[0]) / rel;
  int *g = pw->gbbox;
  pw->nx = 6 + npml + sqrt((g[XHI] - g[XLO]) * (g[XHI] - g[XLO]) + (g[YHI] - g[YLO]) * (g[YHI] - g[YLO]) +
                           (g[ZHI] - g[ZLO]) * (g[ZHI] - g[ZLO]));
  pw->xb = pw->nx - npml;
  pw->Eyi = vh_alloc(s, (size_t)pw->nx + 1, sizeof(real));
  pw->Hzi = vh_alloc(s, (size_t)pw->nx + 1, sizeof(real));
  real **pa[7] = {&pw->Pyi, &pw->PPyi, &pw->Bzi, &pw->adx, &pw->bdx, &pw->ahx, &pw->bhx};
  for (int q = 0; q < 7; q++) *pa[q] = vh_alloc(s, npml, sizeof(real));
  for (int i = 0; i < npml; i++) {
    real depth, sprof;
    depth = abs(i) / (real)npml;
    sprof = 0.5 * dt / s->eps0 * pow(depth, 3.77) * 4.0 / s->eta0 / d[0];
    pw->bdx[i] = 1.0 / (1.0 + sprof);
    pw->adx[i] = (1.0 - sprof) / (1.0 + sprof);
    depth = (abs(i) + 0.5) / (real)(npml);
    sprof = 0.5 * dt / s->eps0 * pow(depth, 3.77) * 4.0 / s->eta0 / d[0];
    pw->bhx[i] = 1.0 / (1.0 + sprof);
    pw->ahx[i] = (1.0 - sprof) / (1.0 + sprof);
  }
}

static void init_planewaves(vh_solve *s)
{
  if (s->numPws == 0) return;
  s->useAuxGrid = s->gridType == GT_CUBIC;
  for (int n = 0; n < s->numPws; n++) {
    vh_planewave *pw = &s->pws[n];
    vh_offset_bbox(pw->gbbox, pw->mbbox, s->gibox);
    calc_incident_vectors(s, pw);
    real r0[3];
    r0[0] = s->xlines[(int)pw->ijk0[0] - s->gibox[XLO]];
    r0[1] = s->ylines[(int)pw->ijk0[1] - s->gibox[YLO]];
    r0[2] = s->zlines[(int)pw->ijk0[2] - s->gibox[ZLO]];
    for (int a = 0; a < 3; a++) pw->r0[a] = r0[a];
    switch (s->gridType) {
    case GT_CUBIC: init_aux_grid(s, pw); break;
    case GT_UNIFORM: pw->phaseVelocity = vh_numerical_phase_velocity(s, deg2rad(s, pw->theta), deg2rad(s, pw->phi)); break;
    default: pw->phaseVelocity = s->c0; break;
    }
    bool edge[6], inc[6];
    for (int f = 0; f < 6; f++) {
      edge[f] = pw->isActive[f];
      int t = vh_outer_type(s, f);
      if (pw->gbbox[f] == s->gibox[f] && (t == BT_PMC || t == BT_PERIODIC)) edge[f] = true;
    }
    for (int face = 0; face < 6; face++) {
      int a = face >> 1, bb[6];
      for (int q = 0; q < 6; q++) bb[q] = pw->gbbox[q];
      /* One cell straddling the TF/SF face: [g-1, g] on low faces, [g, g+1] on high. */
      if (face & 1) { bb[2 * a] = pw->gbbox[face]; bb[2 * a + 1] = pw->gbbox[face] + 1; }
      else { bb[2 * a] = pw->gbbox[face] - 1; bb[2 * a + 1] = pw->gbbox[face]; }
      for (int q = 0; q < 6; q++) inc[q] = edge[q];
      inc[2 * a] = (face & 1) ? true : false;
      inc[2 * a + 1] = (face & 1) ? false : true;
      vh_set_field_limits(bb, pw->flim[face], inc);
    }
  }
}

/* ---------------------------------------------------------------------- */
/* Observers (observer.c:398-520).                                         */
/* ---------------------------------------------------------------------- */

static void init_observers(vh_solve *s)
{
  real dt = (float)s->dt;
  s->obsDt = dt;
  if (!s->isOT) {
    s->startTimeStep = 0UL;
    s->stopTimeStep = s->numTimeSteps - 1UL;
  }
  s->startTime = s->startTimeStep * dt;
  s->stopTime = s->stopTimeStep * dt;
  s->numOutTimeSteps = s->stopTimeStep - s->startTimeStep + 1UL;
  if (!s->isOF) {
    s->startFreq = 0.0;
    s->numFreq = fmax(s->numTimeSteps / 10, 1);
    s->stepFreq = 1.0 / (s->numTimeSteps * s->dt);
    s->stopFreq = s->startFreq + (s->numFreq - 1) * s->stepFreq;
  } else {
    s->stepFreq = (s->stopFreq - s->startFreq) / (s->numFreq - 1);
  }
  int bbox[6] = {0}, step[3] = {0};
  int nwf = s->numWaveforms;
  int *wfFreqObs = vh_alloc(s, nwf, sizeof(int));
  for (int w = 0; w < nwf; w++) {
    vh_add_observer(s, bbox, step, s->waveforms[w].name, OF_ASCII, OD_TIME, OQ_WF, true, w);
    wfFreqObs[w] = vh_add_observer(s, bbox, step, s->waveforms[w].name, OF_ASCII, OD_FREQ, OQ_WF, true, w);
  }
  for (int n = 0; n < s->numObs; n++) {
    vh_observer *o = &s->obs[n];
    vh_offset_bbox(o->gbbox, o->mbbox, s->gibox);
    o->numComp = o->quantity == OQ_WF ? 1 : (o->quantity == OQ_EH ? 6 : 3);
    if (o->format == OF_ASCII && o->domain == OD_FREQ) {
      o->dft_re = vh_alloc(s, (size_t)o->numComp * s->numFreq, sizeof(real));
      o->dft_im = vh_alloc(s, (size_t)o->numComp * s->numFreq, sizeof(real));
    }
    if (o->domain == OD_FREQ && o->quantity != OQ_WF) o->waveformObserver = wfFreqObs[o->waveform];
  }
  free(wfFreqObs);
  s->omega = vh_alloc(s, s->numFreq, sizeof(real));
  for (unsigned long f = 0; f < s->numFreq; f++)
    s->omega[f] = 2.0 * s->pi * (s->startFreq + f * s->stepFreq);
}

/* ---------------------------------------------------------------------- */

vh_solve *vh_create(void)
{
  vh_solve *s = calloc(1, sizeof(vh_solve));
  if (!s) return NULL;
  s->courantNumber = -1.0;
  /* physical.c:37-48 */
  s->pi = 4.0 * atan(1.0);
  s->c0 = 299792458;
  s->mu0 = 4.0 * s->pi * 1e-7;
  s->eps0 = 1.0 / (s->mu0 * s->c0 * s->c0);
  s->eta0 = sqrt(s->mu0 / s->eps0);
  s->openOutputs = 1;
  return s;
}

const char *vh_error(const vh_solve *s) { return s->err; }

/* Create a directory and its parents (like mkdir -p). */
static void make_dirs(const char *path)
{
  char buf[VH_PATH];
  snprintf(buf, sizeof(buf), "%s", path);
  for (char *p = buf + 1; *p; p++)
    if (*p == '/') {
      *p = '\0';
      mkdir(buf, 0777);
      *p = '/';
    }
  mkdir(buf, 0777);
}

void vh_set_dirs(vh_solve *s, const char *inputDir, const char *outputDir)
{
  if (inputDir) strncpy(s->inputDir, inputDir, VH_PATH - 1);
  if (outputDir && outputDir[0]) {
    strncpy(s->outputDir, outputDir, VH_PATH - 1);
    make_dirs(s->outputDir);
  }
}

int vh_init(vh_solve *s)
{
  jmp_buf jb;
  if (setjmp(jb)) {
    s->jb = NULL;
    return -1;
  }
  s->jb = &jb;
  /* initSimulation (simulation.c:113-125) */
  if (s->courantNumber < 0.0) s->courantNumber = sqrt(3.0) / 2.0;
  if (!s->xlines || !s->ylines || !s->zlines) vh_fail(s, "mesh has no mesh lines");
  vh_expand_voxels(s);
  /* initGrid (grid.c:327-355) */
  init_external_surface_parameters(s);
  set_grid_extents(s);
  init_field_array_limits(s);
  alloc_grid(s);
  init_cell_edges(s);
  set_grid_type(s);
  /* vulture.c:106-138 */
  init_boundaries(s);
  init_media(s);
  {
    int mask = 0x20 | 0x8 | 0x2; /* XLO|YLO|ZLO */
    vh_set_medium_on_grid(s, s->ggbox, MT_FREE_SPACE, (FaceMask)mask);
  }
  init_blocks(s);
  init_lines(s);
  init_internal_surfaces(s);
  init_external_pec(s);
  init_pml(s);
  init_mur(s);
  init_waveforms(s);
  init_sources(s);
  init_planewaves(s);
  init_observers(s);
  s->jb = NULL;
  return 0;
}

static void free_rc(vh_rcfilter *rc)
{
  free(rc->q0); free(rc->q1); free(rc->q2); free(rc->poles); free(rc->residues);
}

void vh_release_grid_arrays(vh_solve *s)
{
  real **many[] = {&s->Ex, &s->Ey, &s->Ez, &s->Hx, &s->Hy, &s->Hz, &s->alphaEx, &s->alphaEy, &s->alphaEz,
                   &s->betaEx, &s->betaEy, &s->betaEz, &s->gammaHx, &s->gammaHy, &s->gammaHz};
  for (size_t q = 0; q < sizeof(many) / sizeof(many[0]); q++) {
    free(*many[q]);
    *many[q] = NULL;
  }
  for (int r = 0; r < 6; r++) {
    real **aux[] = {&s->PPx[r], &s->PPy[r], &s->PPz[r], &s->Px[r], &s->Py[r], &s->Pz[r], &s->Bx[r], &s->By[r],
                    &s->Bz[r]};
    for (size_t q = 0; q < sizeof(aux) / sizeof(aux[0]); q++) {
      free(*aux[q]);
      *aux[q] = NULL;
    }
  }
  for (int d = 0; d < s->numDebye; d++)
    for (int f = 0; f < 3; f++) {
      free(s->debye[d].J[f]);
      free(s->debye[d].lastE[f]);
      s->debye[d].J[f] = NULL;
      s->debye[d].lastE[f] = NULL;
    }
}

void vh_destroy(vh_solve *s)
{
  if (!s) return;
  vh_close_observers(s);
  real *many[] = {s->xlines, s->ylines, s->zlines, s->dex, s->dey, s->dez, s->dhx, s->dhy, s->dhz,
                  s->idex, s->idey, s->idez, s->idhx, s->idhy, s->idhz, s->Ex, s->Ey, s->Ez, s->Hx,
                  s->Hy, s->Hz, s->alphaEx, s->alphaEy, s->alphaEz, s->betaEx, s->betaEy, s->betaEz,
                  s->gammaHx, s->gammaHy, s->gammaHz, s->adx, s->bdx, s->gdx, s->kdx, s->ady, s->bdy,
                  s->gdy, s->kdy, s->adz, s->bdz, s->gdz, s->kdz, s->ahx, s->bhx, s->ghx, s->khx, s->ahy,
                  s->bhy, s->ghy, s->khy, s->ahz, s->bhz, s->ghz, s->khz, s->ibdx, s->ibdy, s->ibdz,
                  s->ibhx, s->ibhy, s->ibhz, s->omega};
  for (size_t q = 0; q < sizeof(many) / sizeof(many[0]); q++) free(many[q]);
  for (int r = 0; r < 6; r++) {
    free(s->PPx[r]); free(s->PPy[r]); free(s->PPz[r]); free(s->Px[r]); free(s->Py[r]);
    free(s->Pz[r]); free(s->Bx[r]); free(s->By[r]); free(s->Bz[r]);
  }
  for (int m = 0; m < s->numMedia; m++) {
    free(s->media[m].residues); free(s->media[m].poles); free(s->media[m].dalpha); free(s->media[m].dbeta);
  }
  free(s->media);
  for (int b = 0; b < s->numBounds; b++)
    if (s->bounds[b].hasRc)
      for (int f = 0; f < 16; f++) free_rc(&s->bounds[b].rc[f]);
  free(s->bounds);
  free(s->blocks);
  vh_free_voxels(s);
  free(s->surfaces);
  free(s->wireNames);
  free(s->wireTypes);
  free(s->lines);
  for (int d = 0; d < s->numDebye; d++)
    for (int f = 0; f < 3; f++) { free(s->debye[d].J[f]); free(s->debye[d].lastE[f]); }
  free(s->debye);
  for (int n = 0; n < s->numSibc; n++) {
    free(s->sibc[n].zeta); free(s->sibc[n].old); free(s->sibc[n].Etan);
    free(s->sibc[n].isAdjA); free(s->sibc[n].isAdjB);
  }
  free(s->sibc);
  for (int w = 0; w < s->numWaveforms; w++) free(s->waveforms[w].table);
  free(s->waveforms);
  free(s->sources);
  for (int p = 0; p < s->numPws; p++) {
    vh_planewave *pw = &s->pws[p];
    free(pw->Eyi); free(pw->Hzi); free(pw->Pyi); free(pw->PPyi); free(pw->Bzi);
    free(pw->adx); free(pw->bdx); free(pw->ahx); free(pw->bhx);
  }
  free(s->pws);
  for (int o = 0; o < s->numObs; o++) { free(s->obs[o].dft_re); free(s->obs[o].dft_im); }
  free(s->obs);
  free(s);
}
