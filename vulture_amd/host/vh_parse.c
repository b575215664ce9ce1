/*
 * Mesh file reader: the card grammar of the reference (mesh.c:46-191,
 * 269-405) and every card handler the time-stepping core depends on.
 *
 * Cards are grouped in four sections (VM | CE DM BR BE GS | MT BT WT MB TB TW
 * WF EX PW OP FF GE | NT OT OF CN MS XL YL ZL EN).  Handlers receive the line
 * after the two-letter tag, as the reference's do, and scan with the same
 * formats so numbers round exactly as they do there.
 */
#include <ctype.h>
#include "vh_internal.h"

enum { CT_VM = 0, CT_CE, CT_DM, CT_BR, CT_BE, CT_GS, CT_MT, CT_BT, CT_WT, CT_MB, CT_TB, CT_TW,
       CT_WF, CT_EX, CT_PW, CT_OP, CT_FF, CT_GE, CT_NT, CT_OT, CT_OF, CT_CN, CT_MS, CT_XL,
       CT_YL, CT_ZL, CT_EN, CT_NUM };

static const char *CARD[CT_NUM] = {"VM", "CE", "DM", "BR", "BE", "GS", "MT", "BT", "WT", "MB",
                                   "TB", "TW", "WF", "EX", "PW", "OP", "FF", "GE", "NT", "OT",
                                   "OF", "CN", "MS", "XL", "YL", "ZL", "EN"};
/* Section of each card, whether required, whether it may appear once only
 * and whether it closes its section (mesh.c:79-171). */
static const int SECTION[CT_NUM] = {0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2,
                                    3, 3, 3, 3, 3, 3, 3, 3, 3};
static const bool REQUIRED[CT_NUM] = {1, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1,
                                      1, 0, 0, 0, 0, 0, 0, 0, 1};
static const bool SINGLETON[CT_NUM] = {1, 0, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1,
                                       1, 1, 1, 1, 1, 1, 1, 1, 1};
static const bool ENDS_SECTION[CT_NUM] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1,
                                          0, 0, 0, 0, 0, 0, 0, 0, 1};

static const char *FACE_NAME[6] = {"XLO", "XHI", "YLO", "YHI", "ZLO", "ZHI"};

void vh_fail(vh_solve *s, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(s->err, sizeof(s->err), fmt, ap);
  va_end(ap);
  if (s->jb) longjmp(*(jmp_buf *)s->jb, 1);
  fprintf(stderr, "vh: %s\n", s->err);
  abort();
}

void *vh_alloc(vh_solve *s, size_t n, size_t size)
{
  void *p = calloc(n ? n : 1, size);
  if (!p) vh_fail(s, "out of host memory (%zu x %zu bytes)", n, size);
  return p;
}

void vh_grow(vh_solve *s, void **arr, int *cap, int need, size_t elem)
{
  if (need <= *cap) return;
  int nc = *cap ? *cap * 2 : 16;
  while (nc < need) nc *= 2;
  void *p = realloc(*arr, (size_t)nc * elem);
  if (!p) vh_fail(s, "out of host memory");
  memset((char *)p + (size_t)(*cap) * elem, 0, (size_t)(nc - *cap) * elem);
  *arr = p;
  *cap = nc;
}

FILE *vh_open_input(vh_solve *s, const char *name, const char *mode)
{
  FILE *fp = fopen(name, mode);
  if (!fp && name[0] != '/' && s->inputDir[0]) {
    char path[2 * VH_PATH + 2];
    snprintf(path, sizeof(path), "%s/%s", s->inputDir, name);
    fp = fopen(path, mode);
  }
  return fp;
}

FILE *vh_open_output(vh_solve *s, const char *name, const char *mode)
{
  char path[2 * VH_PATH + 2];
  if (s->outputDir[0])
    snprintf(path, sizeof(path), "%s/%s", s->outputDir, name);
  else
    snprintf(path, sizeof(path), "%s", name);
  return fopen(path, mode);
}

/* ---- name lookups and list builders ---- */

int vh_find_medium(const vh_solve *s, const char *name)
{
  for (int m = 0; m < s->numMedia; m++)
    if (strncmp(s->media[m].name, name, VH_TAG) == 0) return m;
  return -1;
}

int vh_find_boundary(const vh_solve *s, const char *name)
{
  for (int b = 0; b < s->numBounds; b++)
    if (strncmp(s->bounds[b].name, name, VH_TAG) == 0) return b;
  return -1;
}

int vh_find_waveform(const vh_solve *s, const char *name)
{
  for (int w = 0; w < s->numWaveforms; w++)
    if (strncmp(s->waveforms[w].name, name, VH_TAG) == 0) return w;
  return -1;
}

int vh_find_wire(const vh_solve *s, const char *name)
{
  for (int w = 0; w < s->numWires; w++)
    if (strncmp(s->wireNames[w], name, VH_TAG) == 0) return w;
  return -1;
}

static int find_source(const vh_solve *s, const char *name)
{
  for (int i = 0; i < s->numSources; i++)
    if (strncmp(s->sources[i].name, name, VH_TAG) == 0) return i;
  return -1;
}

static int find_pw(const vh_solve *s, const char *name)
{
  for (int i = 0; i < s->numPws; i++)
    if (strncmp(s->pws[i].name, name, VH_TAG) == 0) return i;
  return -1;
}

static int find_observer(const vh_solve *s, const char *name)
{
  for (int i = 0; i < s->numObs; i++)
    if (strncmp(s->obs[i].name, name, VH_TAG) == 0) return i;
  return -1;
}

/* medium.c:245-300 */
int vh_add_medium(vh_solve *s, const char *name, int type, real eps_r, real sigma, real mu_r,
                  int numPoles, const double *residues, const double *poles, const char *fileName)
{
  vh_grow(s, (void **)&s->media, &s->capMedia, s->numMedia + 1, sizeof(vh_medium));
  vh_medium *m = &s->media[s->numMedia];
  memset(m, 0, sizeof(*m));
  strncpy(m->name, name, VH_TAG - 1);
  m->type = type;
  m->eps_r = eps_r;
  m->sigma = sigma;
  m->mu_r = mu_r;
  strncpy(m->fileName, fileName ? fileName : "", VH_PATH - 1);
  if (type == MT_DEBYE) {
    if (strlen(m->fileName) == 0) {
      m->numPoles = numPoles;
      m->residues = vh_alloc(s, numPoles, sizeof(double complex));
      m->poles = vh_alloc(s, numPoles, sizeof(double complex));
      for (int p = 0; p < numPoles; p++) {
        m->residues[p] = (double complex)residues[p];
        m->poles[p] = (double complex)poles[p];
      }
    } else {
      /* External Debye parameter file (medium.c:314-377). */
      FILE *fp = vh_open_input(s, m->fileName, "r");
      double t1, t2, t3, t4;
      if (!fp) vh_fail(s, "cannot open Debye parameter file %s", m->fileName);
      if (fscanf(fp, "%d %lf %lf %lf", &m->numPoles, &t1, &t2, &t3) != 4)
        vh_fail(s, "failed to read first line of Debye parameter file %s", m->fileName);
      if (m->numPoles < 0) vh_fail(s, "negative number of poles in %s", m->fileName);
      if (t1 < 1.0) vh_fail(s, "high frequency rel. permittivity in %s must be >=1", m->fileName);
      m->eps_r = (real)t1;
      if (t2 < 0.0) vh_fail(s, "conductivity in %s must be >=0", m->fileName);
      m->sigma = (real)t2;
      if (t3 < 1.0) vh_fail(s, "rel. permeability in %s must be >=1", m->fileName);
      m->mu_r = (real)t3;
      m->residues = vh_alloc(s, m->numPoles, sizeof(double complex));
      m->poles = vh_alloc(s, m->numPoles, sizeof(double complex));
      for (int p = 0; p < m->numPoles; p++) {
        if (fscanf(fp, "%lf %lf %lf %lf", &t1, &t2, &t3, &t4) != 4)
          vh_fail(s, "failed to read parameters from Debye parameter file %s", m->fileName);
        m->residues[p] = t1 + I * t2;
        if (t3 > 0.0) vh_fail(s, "unstable pole in %s", m->fileName);
        m->poles[p] = t3 + I * t4;
      }
      fclose(fp);
    }
  }
  return s->numMedia++;
}

/* boundary.c:222-263 */
int vh_add_boundary(vh_solve *s, const char *name, int type, int numLayers, int order, real n_eff,
                    real refCoeff, real kmax, const char *fileName, real S_TM[2][2], real S_TE[2][2])
{
  vh_grow(s, (void **)&s->bounds, &s->capBounds, s->numBounds + 1, sizeof(vh_boundary));
  vh_boundary *b = &s->bounds[s->numBounds];
  memset(b, 0, sizeof(*b));
  strncpy(b->name, name, VH_TAG - 1);
  b->type = type;
  b->numLayers = numLayers;
  b->order = order;
  b->n_eff = n_eff;
  b->refCoeff = refCoeff;
  b->kmax = kmax;
  strncpy(b->fileName, fileName ? fileName : "", VH_PATH - 1);
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      if (S_TM) b->S_TM[i][j] = S_TM[i][j];
      if (S_TE) b->S_TE[i][j] = S_TE[i][j];
    }
  return s->numBounds++;
}

int vh_add_observer(vh_solve *s, const int mbbox[6], const int step[3], const char *name, int format,
                    int domain, int quantity, bool isInternal, int waveform)
{
  vh_grow(s, (void **)&s->obs, &s->capObs, s->numObs + 1, sizeof(vh_observer));
  vh_observer *o = &s->obs[s->numObs];
  memset(o, 0, sizeof(*o));
  strncpy(o->name, name, VH_TAG - 1);
  o->format = format;
  o->domain = domain;
  o->quantity = quantity;
  o->isInternal = isInternal;
  o->waveform = waveform;
  o->waveformObserver = -1;
  for (int f = 0; f < 6; f++) o->mbbox[f] = mbbox[f];
  for (int d = 0; d < 3; d++) o->step[d] = step[d];
  return s->numObs++;
}

static void add_wire(vh_solve *s, const char *name, int type)
{
  int cap = s->capWires;
  vh_grow(s, (void **)&s->wireNames, &s->capWires, s->numWires + 1, VH_TAG);
  if (s->capWires != cap) {
    int *t = realloc(s->wireTypes, (size_t)s->capWires * sizeof(int));
    if (!t) vh_fail(s, "out of host memory");
    s->wireTypes = t;
  }
  strncpy(s->wireNames[s->numWires], name, VH_TAG - 1);
  s->wireTypes[s->numWires] = type;
  s->numWires++;
}

static bool bbox_is_normal(const int b[6])
{
  /* bounding_box.h:31-34 (the z test compares ZLO with itself). */
  return b[XLO] <= b[XHI] && b[YLO] <= b[YHI] && b[ZLO] <= b[ZLO];
}

static bool bbox_is_within(const int a[6], const int b[6])
{
  return a[XLO] >= b[XLO] && a[XHI] <= b[XHI] && a[YLO] >= b[YLO] && a[YHI] <= b[YHI] &&
         a[ZLO] >= b[ZLO] && a[ZHI] <= b[ZHI];
}

static FaceMask face_mask_from_string(const char *str)
{
  char *junk = NULL;
  long l = strtol(str, &junk, 2);
  if (l < 0 || l > 63L || *junk) return FACE_MASK_ERROR;
  return (FaceMask)l;
}

/* ---- card handlers ---- */

static bool parse_VM(vh_solve *s, char *line)
{
  return sscanf(line, "%d.%d.%d", &s->meshVersion[0], &s->meshVersion[1], &s->meshVersion[2]) == 3;
}

static bool parse_DM(vh_solve *s, char *line)
{
  if (sscanf(line, "%d %d %d", &s->numCells[0], &s->numCells[1], &s->numCells[2]) != 3) return false;
  for (int d = 0; d < 3; d++) {
    if (s->numCells[d] < 1) return false;
    s->numLines[d] = s->numCells[d] + 1;
  }
  s->mbox[XLO] = 0; s->mbox[XHI] = s->numCells[0];
  s->mbox[YLO] = 0; s->mbox[YHI] = s->numCells[1];
  s->mbox[ZLO] = 0; s->mbox[ZHI] = s->numCells[2];
  return true;
}

/* boundary.c:74-104 */
static bool parse_BE(vh_solve *s, char *line)
{
  int depth[6];
  if (sscanf(line, "%d %d %d %d %d %d", &depth[0], &depth[1], &depth[2], &depth[3], &depth[4],
             &depth[5]) != 6)
    return false;
  for (int f = 0; f < 6; f++) {
    if (depth[f] < 0) return false;
    int b = vh_find_boundary(s, FACE_NAME[f]);
    if (b < 0) b = 0;
    s->bounds[b].numLayers = depth[f];
  }
  return true;
}

static void pml_defaults(int *numLayers, int *order, real *n_eff, real *refCoeff, real *kmax)
{
  /* pml.c:1326-1337 */
  *numLayers = 6;
  *order = 4;
  *n_eff = 1.0;
  *refCoeff = -1.0;
  *kmax = 1.0;
}

/* surface.c:140-190 (obsolete BR card) */
static bool parse_BR(vh_solve *s, char *line)
{
  int t[6];
  if (sscanf(line, "%d %d %d %d %d %d", &t[0], &t[1], &t[2], &t[3], &t[4], &t[5]) != 6) return false;
  for (int f = 0; f < 6; f++) {
    int type, numLayers = 0, order = 0;
    real n_eff = 0.0, refCoeff = 0.0, kmax = 0.0;
    switch (t[f]) {
    case -1: type = BT_PEC; refCoeff = -1.0; break;
    case 0: type = BT_PML; pml_defaults(&numLayers, &order, &n_eff, &refCoeff, &kmax); break;
    case 1: type = BT_PMC; refCoeff = 1.0; break;
    default: return false;
    }
    vh_add_boundary(s, FACE_NAME[f], type, numLayers, order, n_eff, refCoeff, kmax, "", NULL, NULL);
  }
  return true;
}

/* boundary.c:107-219 */
static bool parse_BT(vh_solve *s, char *line)
{
  static const char *TYPES[7] = {"PMC", "PML", "PEC", "FREE_SPACE", "PERIODIC", "MUR", "SIBC"};
  char name[VH_TAG] = "", typeStr[VH_TAG] = "", fileName[VH_PATH] = "";
  int type = BT_PEC, numLayers = 0, order = 0;
  real n_eff = 0.0, refCoeff = 0.0, kmax = 0.0;
  real S_TM[2][2] = {{-1.0, 0.0}, {0.0, -1.0}};
  real S_TE[2][2] = {{-1.0, 0.0}, {0.0, -1.0}};
  bool found = false;
  if (sscanf(line, "%31s %31s", name, typeStr) < 2) return false;
  if (vh_find_boundary(s, name) >= 0) return false;
  for (int t = 0; t < 7; t++)
    if (strncmp(typeStr, TYPES[t], VH_TAG) == 0) { type = t; found = true; }
  if (!found) return false;
  switch (type) {
  case BT_PEC: refCoeff = -1.0; break;
  case BT_PMC: refCoeff = 1.0; break;
  case BT_FREE_SPACE: refCoeff = 0.0; break;
  case BT_PERIODIC: refCoeff = 1.0; break;
  case BT_MUR: refCoeff = 0.0; break;
  case BT_PML:
    pml_defaults(&numLayers, &order, &n_eff, &refCoeff, &kmax);
    sscanf(line, "%31s %31s %d %d " VH_RFMT " " VH_RFMT " " VH_RFMT, name, typeStr, &numLayers, &order,
           &n_eff, &refCoeff, &kmax);
    break;
  case BT_SIBC: {
    int n = sscanf(line, "%31s %31s " VH_RFMT " " VH_RFMT " " VH_RFMT " " VH_RFMT " " VH_RFMT " " VH_RFMT
                   " " VH_RFMT " " VH_RFMT, name, typeStr, &S_TM[0][0], &S_TM[0][1], &S_TM[1][0],
                   &S_TM[1][1], &S_TE[0][0], &S_TE[0][1], &S_TE[1][0], &S_TE[1][1]);
    if (n == 10) {
      /* anisotropic */
    } else if (n == 6) {
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) S_TE[i][j] = S_TM[i][j];
    } else {
      if (sscanf(line, "%31s %31s \"%1023[^\"]\"", name, typeStr, fileName) != 3) return false;
    }
    break;
  }
  default: break;
  }
  vh_add_boundary(s, name, type, numLayers, order, n_eff, refCoeff, kmax, fileName, S_TM, S_TE);
  return true;
}

/* medium.c:75-184 */
static bool parse_MT(vh_solve *s, char *line)
{
  static const char *TYPES[4] = {"FREE_SPACE", "PEC", "SIMPLE", "DEBYE"};
  char name[VH_TAG] = "", typeStr[VH_TAG] = "", fileName[VH_PATH] = "";
  double eps_r = 1.0, sigma = 0.0, mu_r = 1.0;
  double residues[3] = {0.0}, poles[3] = {0.0};
  int numPoles = 0, type = MT_FREE_SPACE, n;
  bool found = false;
  if (sscanf(line, "%31s %31s", name, typeStr) < 2) return false;
  if (vh_find_medium(s, name) >= 0 || vh_find_voxel(s, name) >= 0) return false;
  if (strncmp(typeStr, "VOXEL", VH_TAG) == 0) {
    /* voxel object loader (doc/ToDo.md "Voxel object loader"; not in the reference's parser) */
    if (sscanf(line, "%31s %31s \"%1023[^\"]\"", name, typeStr, fileName) != 3 &&
        sscanf(line, "%31s %31s %1023s", name, typeStr, fileName) != 3)
      return false;
    vh_load_voxel(s, name, fileName);
    return true;
  }
  for (int t = 0; t < 4; t++)
    if (strncmp(typeStr, TYPES[t], VH_TAG) == 0) { type = t; found = true; }
  if (!found) return false;
  switch (type) {
  case MT_SIMPLE:
    n = sscanf(line, "%31s %31s %lf %lf %lf", name, typeStr, &eps_r, &sigma, &mu_r);
    if (n >= 3 && eps_r < 1.0) return false;
    if (n >= 4 && sigma < 0.0) return false;
    if (n >= 5 && mu_r < 1.0) return false;
    break;
  case MT_DEBYE:
    n = sscanf(line, "%31s %31s %lf %lf %lf %lf %lf %lf %lf %lf %lf", name, typeStr, &eps_r, &sigma,
               &mu_r, &residues[0], &poles[0], &residues[1], &poles[1], &residues[2], &poles[2]);
    if (n >= 3 && eps_r < 1.0) return false;
    if (n >= 4 && sigma < 0.0) return false;
    if (n >= 5 && mu_r < 1.0) return false;
    if (n == 7 || n == 9 || n == 11) {
      numPoles = (n - 5) / 2;
    } else {
      if (sscanf(line, "%31s %31s \"%1023[^\"]\"", name, typeStr, fileName) != 3) return false;
    }
    break;
  default: break;
  }
  vh_add_medium(s, name, type, (real)eps_r, (real)sigma, (real)mu_r, numPoles, residues, poles, fileName);
  return true;
}

/* wire.c:78-131 */
static bool parse_WT(vh_solve *s, char *line)
{
  char name[VH_TAG] = "", typeStr[VH_TAG] = "";
  if (sscanf(line, "%31s %31s", name, typeStr) < 2) return false;
  if (vh_find_wire(s, name) >= 0) return false;
  if (strncmp(typeStr, "PEC", VH_TAG) == 0) {
    add_wire(s, name, 0);
    return true;
  }
  return false;
}

/* block.c:63-117 */
static bool parse_MB(vh_solve *s, char *line)
{
  char medium[VH_TAG] = "", maskStr[VH_TAG] = "111111";
  int b[6];
  int n = sscanf(line, "%d %d %d %d %d %d %31s %31s", &b[0], &b[1], &b[2], &b[3], &b[4], &b[5], medium,
                 maskStr);
  if (n < 7) return false;
  int m;
  FaceMask mask = FACE_MASK_ALL;
  int v = vh_find_voxel(s, medium);
  if (v >= 0) {
    /* a voxel object placed with its reference point on a node, two optional angles */
    double ang[2] = {0.0, 0.0};
    char extra[VH_TAG];
    int na = sscanf(line, "%*d %*d %*d %*d %*d %*d %*s %lf %lf %31s", &ang[0], &ang[1], extra);
    if (na == 3 || !isfinite(ang[0]) || !isfinite(ang[1])) return false;
    if (na <= 0 && n >= 8) return false;
    if (!bbox_is_within(b, s->mbox) || vh_bbox_type(b) != BB_POINT) return false;
    int node[3] = {b[XLO], b[YLO], b[ZLO]};
    m = vh_place_voxel(s, v, node, ang[0], ang[1]);
    mask = FACE_MASK_ALL;
  } else {
    if (!bbox_is_normal(b) || !bbox_is_within(b, s->mbox) || vh_bbox_type(b) != BB_VOLUME) return false;
    m = vh_find_medium(s, medium);
    if (m < 0) return false;
    if (n <= 8) {
      mask = face_mask_from_string(maskStr);
      if (mask == FACE_MASK_ERROR) return false;
    }
  }
  vh_grow(s, (void **)&s->blocks, &s->capBlocks, s->numBlocks + 1, sizeof(vh_block));
  vh_block *blk = &s->blocks[s->numBlocks++];
  memcpy(blk->mbbox, b, sizeof(b));
  blk->medium = m;
  blk->mask = mask;
  return true;
}

/* surface.c:193-250 */
static bool parse_TB(vh_solve *s, char *line)
{
  char bname[VH_TAG] = "";
  int b[6], orient = 1;
  double angle = 0.0;
  int n = sscanf(line, "%d %d %d %d %d %d %31s %d %lf", &b[0], &b[1], &b[2], &b[3], &b[4], &b[5], bname,
                 &orient, &angle);
  if (n < 7) return false;
  if (!bbox_is_normal(b) || !bbox_is_within(b, s->mbox) || vh_bbox_type(b) != BB_SURFACE) return false;
  int bn = vh_find_boundary(s, bname);
  if (bn < 0) return false;
  if (orient != 1 && orient != -1) return false;
  if (angle < -180.0 || angle > 180.0) return false;
  vh_grow(s, (void **)&s->surfaces, &s->capSurfaces, s->numSurfaces + 1, sizeof(vh_surface));
  vh_surface *sf = &s->surfaces[s->numSurfaces++];
  memcpy(sf->mbbox, b, sizeof(b));
  sf->boundary = bn;
  sf->orient = orient;
  sf->angle = (real)angle;
  return true;
}

/* line.c:92-160 */
static bool parse_TW(vh_solve *s, char *line)
{
  char wname[VH_TAG] = "", lo[VH_TAG] = "", hi[VH_TAG] = "";
  int b[6];
  int n = sscanf(line, "%d %d %d %d %d %d %31s %31s %31s", &b[0], &b[1], &b[2], &b[3], &b[4], &b[5],
                 wname, lo, hi);
  if (n < 7) return false;
  if (!bbox_is_normal(b) || !bbox_is_within(b, s->mbox) || vh_bbox_type(b) != BB_LINE) return false;
  int w = vh_find_wire(s, wname);
  if (w < 0) return false;
  static const char *ENDS[3] = {"THRU", "END", "CORNER"};
  if (n >= 8) {
    int ok = 0;
    for (int e = 0; e < 3; e++) ok |= strncmp(ENDS[e], lo, VH_TAG) == 0;
    if (!ok) vh_fail(s, "invalid wire low end type %s in TW card", lo);
  }
  if (n >= 9) {
    int ok = 0;
    for (int e = 0; e < 3; e++) ok |= strncmp(ENDS[e], hi, VH_TAG) == 0;
    if (!ok) vh_fail(s, "invalid wire high end type %s in TW card", hi);
  }
  vh_grow(s, (void **)&s->lines, &s->capLinesTW, s->numLinesTW + 1, sizeof(vh_line));
  vh_line *l = &s->lines[s->numLinesTW++];
  memcpy(l->mbbox, b, sizeof(b));
  l->wire = w;
  return true;
}

/* waveform.c:168-252 */
static bool parse_WF(vh_solve *s, char *line)
{
  static const char *TYPES[10] = {"GAUSSIAN_PULSE", "NARROW_GAUSSIAN_PULSE", "DIFF_GAUSSIAN_PULSE",
                                  "RICKER_WAVELET", "MOD_GAUSSIAN_PULSE", "COMPACT_PULSE",
                                  "DIFF_COMPACT_PULSE", "MOD_COMPACT_PULSE", "RAMPED_SINUSOID",
                                  "EXTERNAL"};
  char name[VH_TAG] = "", typeStr[VH_TAG] = "", fileName[VH_PATH] = "";
  double size = -1.0, delay = -1.0, width = -1.0, frequency = -1.0;
  int type = WT_GAUSSIAN_PULSE, n;
  bool found = false;
  if (sscanf(line, "%31s %31s", name, typeStr) < 2) return false;
  if (vh_find_waveform(s, name) >= 0) return false;
  for (int t = 0; t < 10; t++)
    if (strncmp(typeStr, TYPES[t], VH_TAG) == 0) { type = t; found = true; }
  if (!found) return false;
  if (type == WT_EXTERNAL) {
    n = sscanf(line, "%31s %31s \"%1023[^\"]\" %lf %lf", name, typeStr, fileName, &size, &delay);
    if (n < 3) return false;
    if (n >= 4 && size < 0.0) return false;
  } else {
    n = sscanf(line, "%31s %31s %lf %lf %lf %lf", name, typeStr, &size, &delay, &width, &frequency);
    if (n >= 3 && size < 0.0) return false;
    if (n >= 5 && width <= 0.0) return false;
    if (n == 6 && frequency <= 0.0) return false;
  }
  vh_grow(s, (void **)&s->waveforms, &s->capWaveforms, s->numWaveforms + 1, sizeof(vh_waveform));
  vh_waveform *w = &s->waveforms[s->numWaveforms++];
  memset(w, 0, sizeof(*w));
  strncpy(w->name, name, VH_TAG - 1);
  w->type = type;
  w->size = (real)size;
  w->width = (real)width;
  w->delay = (real)delay;
  w->frequency = (real)((real)frequency);
  strncpy(w->fileName, fileName, VH_PATH - 1);
  return true;
}

/* Source tag tables (source.c:97-280): 84 tags in groups of three (x, y, z). */
typedef struct { const char *tag; int type; int field; int normal; bool soft; } src_tag;

static bool lookup_source_tag(const char *tag, int *type, int *pol, int *normal, int *field, bool *soft)
{
  /* Each row: tag stem, type, first field, normals for x/y/z (or -1), softness. */
  static const struct { const char *stem; int type; int f0; int nrm[3]; bool soft; } G[] = {
    {"E", ST_EFIELD, EX, {-1, -1, -1}, true},
    {"H", ST_HFIELD, HX, {-1, -1, -1}, true},
    {"=E", ST_EFIELD, EX, {-1, -1, -1}, false},
    {"=H", ST_HFIELD, HX, {-1, -1, -1}, false},
    {"J", ST_ELEC_CURR_DENSITY, EX, {-1, -1, -1}, true},
    {"JM", ST_MAGN_CURR_DENSITY, HX, {-1, -1, -1}, true},
    {"=J", ST_ELEC_CURR_DENSITY, EX, {-1, -1, -1}, false},
    {"=JM", ST_MAGN_CURR_DENSITY, HX, {-1, -1, -1}, false},
    {"I", ST_ELEC_CURRENT, EX, {-1, -1, -1}, true},
    {"IM", ST_MAGN_CURRENT, HX, {-1, -1, -1}, true},
    {"=I", ST_ELEC_CURRENT, EX, {-1, -1, -1}, false},
    {"=IM", ST_MAGN_CURRENT, HX, {-1, -1, -1}, false},
    {"ID", ST_ELEC_CURRENT_MOMENT, EX, {-1, -1, -1}, true},
    {"IMD", ST_MAGN_CURRENT_MOMENT, HX, {-1, -1, -1}, true},
    {"=ID", ST_ELEC_CURRENT_MOMENT, EX, {-1, -1, -1}, false},
    {"=IMD", ST_MAGN_CURRENT_MOMENT, HX, {-1, -1, -1}, false},
    {"V", ST_VOLTAGE, EX, {-1, -1, -1}, true},
    {"=V", ST_VOLTAGE, EX, {-1, -1, -1}, false},
    {"VR", ST_THEVENIN_VOLTAGE, EX, {-1, -1, -1}, true},
    {"IG", ST_NORTON_CURRENT, EX, {-1, -1, -1}, true},
  };
  static const char AX[3] = {'X', 'Y', 'Z'};
  char buf[VH_TAG];
  for (size_t g = 0; g < sizeof(G) / sizeof(G[0]); g++)
    for (int a = 0; a < 3; a++) {
      snprintf(buf, sizeof(buf), "%s%c", G[g].stem, AX[a]);
      if (strncmp(tag, buf, VH_TAG) == 0) {
        *type = G[g].type; *pol = a; *normal = CA_UNDEFINED; *field = G[g].f0 + a; *soft = G[g].soft;
        return true;
      }
    }
  /* Surface current densities JS<pol><normal>, JMS.., =JS.., =JMS.. */
  static const struct { const char *stem; int type; int f0; bool soft; } S[] = {
    {"JS", ST_ELEC_SURF_CURR_DENSITY, EX, true},
    {"JMS", ST_MAGN_SURF_CURR_DENSITY, HX, true},
    {"=JS", ST_ELEC_SURF_CURR_DENSITY, EX, false},
    {"=JMS", ST_MAGN_SURF_CURR_DENSITY, HX, false},
  };
  /* Order within each group of six: XY YZ ZX XZ YX ZY. */
  static const int PN[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 2}, {1, 0}, {2, 1}};
  for (size_t g = 0; g < sizeof(S) / sizeof(S[0]); g++)
    for (int q = 0; q < 6; q++) {
      snprintf(buf, sizeof(buf), "%s%c%c", S[g].stem, AX[PN[q][0]], AX[PN[q][1]]);
      if (strncmp(tag, buf, VH_TAG) == 0) {
        *type = S[g].type; *pol = PN[q][0]; *normal = PN[q][1]; *field = S[g].f0 + PN[q][0];
        *soft = S[g].soft;
        return true;
      }
    }
  return false;
}

/* source.c:350-463 */
static bool parse_EX(vh_solve *s, char *line)
{
  char name[VH_TAG] = "", typeStr[VH_TAG] = "", wfName[VH_TAG] = "";
  int b[6];
  double size = 1.0, delay = 0.0, resist = 0.0;
  int type = ST_EFIELD, pol = XDIR, normal = CA_UNDEFINED, field = EX;
  bool soft = true;
  if (sscanf(line, "%d %d %d %d %d %d %31s %31s %31s", &b[0], &b[1], &b[2], &b[3], &b[4], &b[5], name,
             typeStr, wfName) < 9)
    return false;
  if (find_source(s, name) >= 0) return false;
  if (!lookup_source_tag(typeStr, &type, &pol, &normal, &field, &soft)) return false;
  if (!bbox_is_normal(b) || !bbox_is_within(b, s->mbox)) return false;
  if (type == ST_THEVENIN_VOLTAGE || type == ST_NORTON_CURRENT) {
    sscanf(line, "%d %d %d %d %d %d %31s %31s %31s %lf %lf %lf", &b[0], &b[1], &b[2], &b[3], &b[4],
           &b[5], name, typeStr, wfName, &resist, &size, &delay);
    if (vh_bbox_type(b) != BB_LINE) return false;
    char rname[VH_TAG];
    snprintf(rname, sizeof(rname), "__VR_RS_%d__", s->numSources + 1);
    vh_add_medium(s, rname, MT_SIMPLE, 1.0, 0.0, 1.0, 0, NULL, NULL, "");
  } else {
    sscanf(line, "%d %d %d %d %d %d %31s %31s %31s %lf %lf", &b[0], &b[1], &b[2], &b[3], &b[4], &b[5],
           name, typeStr, wfName, &size, &delay);
  }
  int w = vh_find_waveform(s, wfName);
  if (w < 0) return false;
  vh_grow(s, (void **)&s->sources, &s->capSources, s->numSources + 1, sizeof(vh_source));
  vh_source *src = &s->sources[s->numSources++];
  memset(src, 0, sizeof(*src));
  strncpy(src->name, name, VH_TAG - 1);
  src->type = type; src->pol = pol; src->normal = normal; src->field = field; src->isSoft = soft;
  src->size = (real)size; src->delay = (real)delay; src->resistance = (real)resist;
  src->waveform = w;
  memcpy(src->mbbox, b, sizeof(b));
  return true;
}

/* planewave.c:183-275 */
static bool parse_PW(vh_solve *s, char *line)
{
  char name[VH_TAG] = "", wfName[VH_TAG] = "", maskStr[VH_TAG] = "111111";
  int b[6];
  double size = 1.0, delay = 0.0, theta = 0.0, phi = 0.0, eta = 0.0;
  bool active[6] = {true, true, true, true, true, true};
  int n = sscanf(line, "%d %d %d %d %d %d %31s %31s %lf %lf %lf %31s %lf %lf", &b[0], &b[1], &b[2], &b[3],
                 &b[4], &b[5], name, wfName, &theta, &phi, &eta, maskStr, &size, &delay);
  if (n < 11) return false;
  if (find_pw(s, name) >= 0) return false;
  if (!bbox_is_normal(b) || !bbox_is_within(b, s->mbox)) return false;
  int w = vh_find_waveform(s, wfName);
  if (w < 0) return false;
  if (theta < 0.0 || theta > 180.0) return false;
  if (phi < 0.0 || phi > 360.0) return false;
  if (eta < 0 || eta >= 360.0) return false;
  if (n >= 12) {
    if (strlen(maskStr) != 6) return false;
    for (int f = 0; f < 6; f++) {
      if (maskStr[f] == '0') active[f] = false;
      else if (maskStr[f] == '1') active[f] = true;
      else return false;
    }
  }
  if (n >= 13 && size < 0.0) return false;
  vh_grow(s, (void **)&s->pws, &s->capPws, s->numPws + 1, sizeof(vh_planewave));
  vh_planewave *pw = &s->pws[s->numPws++];
  memset(pw, 0, sizeof(*pw));
  strncpy(pw->name, name, VH_TAG - 1);
  memcpy(pw->mbbox, b, sizeof(b));
  pw->waveform = w;
  for (int f = 0; f < 6; f++) pw->isActive[f] = active[f];
  pw->theta = (real)theta; pw->phi = (real)phi; pw->eta = (real)eta;
  pw->size = (real)size; pw->delay = (real)delay;
  return true;
}

/* observer.c:193-302 */
static bool parse_OP(vh_solve *s, char *line)
{
  static const char *TYPES[3] = {"TDOM_ASCII", "FDOM_ASCII", "TDOM_BINARY"};
  static const int FMT[3] = {OF_ASCII, OF_ASCII, OF_BINARY};
  static const int DOM[3] = {OD_TIME, OD_FREQ, OD_TIME};
  char typeStr[VH_TAG] = "", wfName[VH_TAG] = "", name[VH_TAG] = "";
  int b[6] = {0}, step[3] = {1, 1, 1}, format = OF_ASCII, domain = OD_TIME, wf = 0;
  bool found = false;
  if (sscanf(line, "%d %d %d %d %d %d %31s %31s", &b[0], &b[1], &b[2], &b[3], &b[4], &b[5], name,
             typeStr) < 8)
    return false;
  if (find_observer(s, name) >= 0) return false;
  if (!bbox_is_normal(b) || !bbox_is_within(b, s->mbox)) return false;
  for (int t = 0; t < 3; t++)
    if (strncmp(typeStr, TYPES[t], VH_TAG) == 0) { format = FMT[t]; domain = DOM[t]; found = true; }
  if (!found) return false;
  if (format == OF_ASCII) {
    if (vh_bbox_type(b) != BB_POINT) return false;
    sscanf(line, "%d %d %d %d %d %d %31s %31s %31s", &b[0], &b[1], &b[2], &b[3], &b[4], &b[5], name,
           typeStr, wfName);
  } else {
    sscanf(line, "%d %d %d %d %d %d %31s %31s %d %d %d", &b[0], &b[1], &b[2], &b[3], &b[4], &b[5], name,
           typeStr, &step[0], &step[1], &step[2]);
    if (step[0] < 0 || step[1] < 0 || step[2] < 0) return false;
  }
  if (strncmp(wfName, "", VH_TAG) != 0) {
    wf = vh_find_waveform(s, wfName);
    if (wf < 0) return false;
  }
  vh_add_observer(s, b, step, name, format, domain, OQ_EH, false, wf);
  return true;
}

static bool parse_NT(vh_solve *s, char *line)
{
  unsigned long n = 0UL;
  if (sscanf(line, "%lu", &n) != 1) return false;
  s->numTimeSteps = n;
  return true;
}

static bool parse_OT(vh_solve *s, char *line)
{
  if (sscanf(line, "%lu %lu", &s->startTimeStep, &s->stopTimeStep) != 2) return false;
  if (s->stopTimeStep < s->startTimeStep) return false;
  s->isOT = true;
  return true;
}

static bool parse_OF(vh_solve *s, char *line)
{
  double fstart = -1.0, fstop = -1.0;
  unsigned long num = 0;
  if (sscanf(line, "%lf %lf %lu", &fstart, &fstop, &num) != 3) return false;
  if (fstart < 0) return false;
  s->startFreq = (real)fstart;
  if (fstop < fstart) return false;
  s->stopFreq = (real)fstop;
  if (num < 1UL) return false;
  s->numFreq = num;
  s->isOF = true;
  return true;
}

static bool parse_CN(vh_solve *s, char *line)
{
  double cn;
  if (sscanf(line, "%lf", &cn) != 1) return false;
  if (cn >= 1.0) return false;
  s->courantNumber = (real)cn;
  return true;
}

/* grid.c:190-232 */
static bool parse_MS(vh_solve *s, char *line)
{
  double d[3];
  int n = sscanf(line, "%lf %lf %lf", &d[0], &d[1], &d[2]);
  if (n == 1) d[1] = d[2] = d[0];
  else if (n != 3) return false;
  for (int a = 0; a < 3; a++)
    if (d[a] <= 0.0) return false;
  real **lines[3] = {&s->xlines, &s->ylines, &s->zlines};
  for (int a = 0; a < 3; a++) {
    *lines[a] = vh_alloc(s, s->numLines[a], sizeof(real));
    real del = (real)(d[a]);
    for (int i = 0; i < s->numLines[a]; i++) (*lines[a])[i] = i * del;
  }
  return true;
}

static bool parse_lines(vh_solve *s, FILE *fp, int axis)
{
  real **lines[3] = {&s->xlines, &s->ylines, &s->zlines};
  *lines[axis] = vh_alloc(s, s->numLines[axis], sizeof(real));
  for (int k = 0; k < s->numLines[axis]; k++) {
    double v;
    if (fscanf(fp, "%lf", &v) != 1) return false;
    (*lines[axis])[k] = (real)v;
  }
  return true;
}

bool vh_parse_card(vh_solve *s, int card, char *line, FILE *fp)
{
  switch (card) {
  case CT_VM: return parse_VM(s, line);
  case CT_CE: strncpy(s->comment, line, sizeof(s->comment) - 1); return true;
  case CT_DM: return parse_DM(s, line);
  case CT_BR: return parse_BR(s, line);
  case CT_BE: return parse_BE(s, line);
  case CT_MT: return parse_MT(s, line);
  case CT_BT: return parse_BT(s, line);
  case CT_WT: return parse_WT(s, line);
  case CT_MB: return parse_MB(s, line);
  case CT_TB: return parse_TB(s, line);
  case CT_TW: return parse_TW(s, line);
  case CT_WF: return parse_WF(s, line);
  case CT_EX: return parse_EX(s, line);
  case CT_PW: return parse_PW(s, line);
  case CT_OP: return parse_OP(s, line);
  case CT_FF: return true; /* far-field requests are ignored by the reference too */
  case CT_NT: return parse_NT(s, line);
  case CT_OT: return parse_OT(s, line);
  case CT_OF: return parse_OF(s, line);
  case CT_CN: return parse_CN(s, line);
  case CT_MS: return parse_MS(s, line);
  case CT_XL: return parse_lines(s, fp, 0);
  case CT_YL: return parse_lines(s, fp, 1);
  case CT_ZL: return parse_lines(s, fp, 2);
  default: return true;
  }
}

/* Trim and strip '#' comments (mesh.c:409-470). */
static void clean_line(char *str)
{
  char *c = strchr(str, '#');
  if (c) *c = '\0';
  size_t n = strlen(str);
  while (n > 0 && isspace((unsigned char)str[n - 1])) str[--n] = '\0';
  size_t st = 0;
  while (str[st] && isspace((unsigned char)str[st])) st++;
  if (st) memmove(str, str + st, strlen(str + st) + 1);
}

static int card_type(const char *line)
{
  if (strlen(line) < 2) return -1;
  for (int c = 0; c < CT_NUM; c++)
    if (strncmp(line, CARD[c], 2) == 0) return c;
  return -1;
}

/* Predefined boundaries, wires and media (mesh.c:246-266). */
static void init_mesh(vh_solve *s)
{
  vh_add_boundary(s, "PEC", BT_PEC, 0, 0, 1.0, -1.0, 1.0, "", NULL, NULL);
  vh_add_boundary(s, "PMC", BT_PMC, 0, 0, 1.0, +1.0, 1.0, "", NULL, NULL);
  vh_add_boundary(s, "FREE_SPACE", BT_FREE_SPACE, 0, 0, 1.0, 0.0, 1.0, "", NULL, NULL);
  add_wire(s, "PEC", 0);
  add_wire(s, "FREE_SPACE", 1);
  vh_add_medium(s, "FREE_SPACE", MT_SIMPLE, 1.0, 0.0, 1.0, 0, NULL, NULL, "");
  vh_add_medium(s, "PEC", MT_PEC, 1.0, 1e8, 1.0, 0, NULL, NULL, "");
}

int vh_read_mesh(vh_solve *s, const char *path)
{
  jmp_buf jb;
  FILE *fp = NULL;
  if (setjmp(jb)) {
    if (fp) fclose(fp);
    s->jb = NULL;
    return -1;
  }
  s->jb = &jb;
  if (!s->inputDir[0]) {
    const char *sl = strrchr(path, '/');
    if (sl) {
      size_t n = (size_t)(sl - path);
      if (n >= VH_PATH) n = VH_PATH - 1;
      memcpy(s->inputDir, path, n);
      s->inputDir[n] = '\0';
    }
  }
  init_mesh(s);
  fp = fopen(path, "r");
  if (!fp) vh_fail(s, "cannot open input mesh file %s", path);
  char raw[1024], line[1024];
  unsigned long lineNumber = 0;
  int section = 0;
  while (fgets(raw, sizeof(raw) - 1, fp)) {
    lineNumber++;
    strncpy(line, raw, sizeof(line));
    clean_line(line);
    if (!line[0]) continue;
    int c = card_type(line);
    if (c < 0) vh_fail(s, "invalid card on line %lu: %s", lineNumber, raw);
    if (SECTION[c] != section)
      vh_fail(s, "card type %s on line %lu is invalid in mesh section %d", CARD[c], lineNumber, section);
    if (s->isFound[c] && SINGLETON[c])
      vh_fail(s, "card type %s on line %lu has already been found", CARD[c], lineNumber);
    s->isFound[c] = 1;
    if (!vh_parse_card(s, c, line + 2, fp))
      vh_fail(s, "failed to parse %s card on line %lu: %s", CARD[c], lineNumber, raw);
    switch (c) {
    case CT_EX: case CT_PW: s->foundSource = true; break;
    case CT_OP: case CT_FF: s->foundOutput = true; break;
    case CT_GE:
      if (!s->foundSource) vh_fail(s, "no source found in section 2 of mesh file");
      break;
    case CT_MS: s->isFound[CT_MS] = s->isFound[CT_XL] = s->isFound[CT_YL] = s->isFound[CT_ZL] = 1; break;
    case CT_XL: s->isFound[CT_MS] = s->isFound[CT_XL] = 1; break;
    case CT_YL: s->isFound[CT_MS] = s->isFound[CT_YL] = 1; break;
    case CT_ZL: s->isFound[CT_MS] = s->isFound[CT_ZL] = 1; break;
    case CT_EN:
      if (!s->isFound[CT_MS] || !(s->isFound[CT_XL] && s->isFound[CT_YL] && s->isFound[CT_ZL]))
        vh_fail(s, "no mesh lines found in section 3 of mesh file");
      break;
    default: break;
    }
    if (ENDS_SECTION[c]) {
      for (int k = 0; k < CT_NUM; k++)
        if (SECTION[k] == section && REQUIRED[k] && !s->isFound[k])
          vh_fail(s, "required card %s not found in section %d", CARD[k], section);
      section++;
    }
  }
  fclose(fp);
  fp = NULL;
  if (section < 4) vh_fail(s, "mesh file ended in section %d", section);
  s->jb = NULL;
  return 0;
}
