/*
 * vh_voxel.c - voxel objects: materials from a voxel array instead of one
 * MB card per run of cells.
 *
 * The reference plans this loader and does not implement it (doc/ToDo.md,
 * "Voxel object loader": a .vox header naming a binary key array, a key table
 * of materials, placement by an MB card at a node with two rotation angles,
 * "method 1": cells are assigned at set-up and become material blocks).  This
 * file implements that plan for every medium type, Debye included.
 *
 * Cards
 *   MT <name> VOXEL "<file.vox>"            declare a voxel object
 *   MB i i j j k k <name> [theta [phi]]     place it: the point bounding box
 *                                           is the node that the object's
 *                                           reference point sits on; angles
 *                                           in degrees
 *
 * .vox file (text, fields separated by white space)
 *   <data file>                name of the key array, quoted or bare; found
 *                              as any other input file
 *   <bytes>                    bytes per key: 1, 2 or 4
 *   <nx> <ny> <nz>             voxels per axis
 *   <dx> <dy> <dz>             voxel size (m)
 *   <xref> <yref> <zref>       reference point (m from the array's low corner)
 *   then one line per key, either
 *   <key> <eps_r> <sigma> <mu_r>   a SIMPLE medium of its own, or
 *   <key> <medium>                 a medium declared by an earlier MT card
 *                                  (any type, DEBYE included)
 * The data file holds nx*ny*nz little-endian unsigned keys, voxel (i, j, k)
 * at ((i*ny)+j)*nz+k, and nothing else.  A key with no line in the table
 * leaves a cell's material as it was.
 *
 * Placement.  A point v of the object (m, from its reference point) lies at
 * P0 + Rz(phi) Ry(theta) v in the mesh, P0 the node of the MB card.  A mesh
 * cell takes the key of the voxel that holds its centre; the cells tried are
 * those whose centres lie in the axis-aligned box around the placed object,
 * cut to the mesh.  Positions are computed in double from the mesh lines;
 * multiples of 90 degrees use exact sines and cosines, so quarter turns of a
 * voxel array on a mesh of its own pitch map voxels to cells exactly.
 *
 * Blocks.  Each run along k of cells with one mapped key becomes one block,
 * in order of i, then j, then k, and the blocks take the MB card's place in
 * the block list (cards after it override them, as for any block).  A block
 * owns the electric edges of its low x and y faces and of its low z face; of
 * its high z face when the next cell along k is not in another run, and never
 * of its high x and y faces (the neighbouring cell's low faces).  So no edge
 * lies in two blocks of one placement, which a Debye medium needs (each edge
 * holds one polarisation current), and a placement of single-cell-wide runs
 * reproduces the face mask 101011 that one MB card per (i, j) run uses.
 */
#include "vh_internal.h"

#include <errno.h>
#include <stdint.h>

#define VOXEL_MASK_LOW 0x2A       /* XLO | YLO | ZLO */
#define VOXEL_MASK_ZHI 0x01

static int find_voxel(const vh_solve *s, const char *name)
{
  for (int v = 0; v < s->numVoxels; v++)
    if (strncmp(s->voxels[v].name, name, VH_TAG) == 0) return v;
  return -1;
}

int vh_find_voxel(const vh_solve *s, const char *name) { return find_voxel(s, name); }

static uint32_t voxel_key(const vh_voxel *o, size_t id)
{
  const unsigned char *p = o->data + id * (size_t)o->bytes;
  uint32_t k = 0;
  for (int b = o->bytes - 1; b >= 0; b--) k = (k << 8) | p[b];
  return k;
}

/* Medium of a key, or -1 when the key has no line in the table. */
static int key_medium(const vh_voxel *o, uint32_t key)
{
  int lo = 0, hi = o->numKeys - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    if (o->keys[mid] == key) return o->keyMedium[mid];
    if (o->keys[mid] < key) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

/* Next white-space separated token of the header, honouring double quotes. */
static bool next_token(FILE *fp, char *buf, size_t cap)
{
  int c;
  do c = fgetc(fp); while (c != EOF && (c == ' ' || c == '\t' || c == '\r' || c == '\n'));
  if (c == EOF) return false;
  size_t n = 0;
  if (c == '"') {
    while ((c = fgetc(fp)) != EOF && c != '"')
      if (n + 1 < cap) buf[n++] = (char)c;
    if (c != '"') return false;
  } else {
    do {
      if (n + 1 < cap) buf[n++] = (char)c;
      c = fgetc(fp);
    } while (c != EOF && c != ' ' && c != '\t' && c != '\r' && c != '\n');
    if (c != EOF) ungetc(c, fp);   /* the rest of the line stays for the caller */
  }
  buf[n] = 0;
  return true;
}

static bool parse_long(const char *t, long *v)
{
  char *end = NULL;
  errno = 0;
  *v = strtol(t, &end, 10);
  return errno == 0 && end != t && *end == 0;
}

static bool parse_double(const char *t, double *v)
{
  char *end = NULL;
  errno = 0;
  *v = strtod(t, &end);
  return errno == 0 && end != t && *end == 0 && isfinite(*v);
}

static int cmp_key(const void *a, const void *b)
{
  uint32_t x = ((const uint32_t *)a)[0], y = ((const uint32_t *)b)[0];
  return x < y ? -1 : x > y;
}

/* Read the header and the key array of MT <name> VOXEL "<file>".  The object
 * is attached to the solve before any check can fail, so vh_destroy frees what
 * was read. */
void vh_load_voxel(vh_solve *s, const char *name, const char *file)
{
  vh_grow(s, (void **)&s->voxels, &s->capVoxels, s->numVoxels + 1, sizeof(vh_voxel));
  vh_voxel *o = &s->voxels[s->numVoxels++];
  memset(o, 0, sizeof(*o));
  strncpy(o->name, name, VH_TAG - 1);

  FILE *fp = vh_open_input(s, file, "r");
  if (!fp) vh_fail(s, "cannot open voxel file %s", file);
  char tok[VH_PATH], dataName[VH_PATH] = "";
  long iv[4];
  double dv[6];
  bool ok = next_token(fp, dataName, sizeof(dataName));
  ok = ok && next_token(fp, tok, sizeof(tok)) && parse_long(tok, &iv[0]);
  for (int a = 0; a < 3 && ok; a++) ok = next_token(fp, tok, sizeof(tok)) && parse_long(tok, &iv[1 + a]);
  for (int a = 0; a < 6 && ok; a++) ok = next_token(fp, tok, sizeof(tok)) && parse_double(tok, &dv[a]);
  if (!ok) {
    fclose(fp);
    vh_fail(s, "voxel file %s: header needs a data file, bytes, nx ny nz, dx dy dz, xref yref zref", file);
  }
  if (iv[0] != 1 && iv[0] != 2 && iv[0] != 4) {
    fclose(fp);
    vh_fail(s, "voxel file %s: %ld bytes per key (1, 2 or 4)", file, iv[0]);
  }
  for (int a = 0; a < 3; a++)
    if (iv[1 + a] < 1 || iv[1 + a] > (1L << 20) || !(dv[a] > 0.0)) {
      fclose(fp);
      vh_fail(s, "voxel file %s: voxel counts must be positive and sizes above zero", file);
    }
  o->bytes = (int)iv[0];
  for (int a = 0; a < 3; a++) { o->n[a] = (int)iv[1 + a]; o->d[a] = dv[a]; o->ref[a] = dv[3 + a]; }

  /* key table: one line per key, two or four fields */
  char line[2 * VH_PATH];
  int c;
  while ((c = fgetc(fp)) != EOF && c != '\n') {}
  while (fgets(line, sizeof(line), fp)) {
    char f[5][VH_PATH];
    int nf = sscanf(line, "%1023s %1023s %1023s %1023s %1023s", f[0], f[1], f[2], f[3], f[4]);
    if (nf <= 0) continue;
    long key;
    bool good = (nf == 2 || nf == 4) && parse_long(f[0], &key) && key >= 0 &&
                (o->bytes == 4 ? key <= 0xFFFFFFFFL : key < (1L << (8 * o->bytes)));
    int m = -1;
    if (good && nf == 2) {
      m = vh_find_medium(s, f[1]);
      good = m >= 0;
    } else if (good) {
      double er, sg, mr;
      good = parse_double(f[1], &er) && parse_double(f[2], &sg) && parse_double(f[3], &mr) && er >= 1.0 &&
             sg >= 0.0 && mr >= 1.0;
      char mname[VH_TAG + 16];
      snprintf(mname, sizeof(mname), "%s#%ld", name, key);
      good = good && strlen(mname) < VH_TAG && vh_find_medium(s, mname) < 0;
      if (good) m = vh_add_medium(s, mname, MT_SIMPLE, (real)er, (real)sg, (real)mr, 0, NULL, NULL, "");
    }
    if (!good) {
      fclose(fp);
      vh_fail(s, "voxel file %s: bad key line: %s", file, line);
    }
    o = &s->voxels[s->numVoxels - 1];
    o->keys = realloc(o->keys, (size_t)(o->numKeys + 1) * 2 * sizeof(uint32_t));
    if (!o->keys) {
      fclose(fp);
      vh_fail(s, "out of host memory");
    }
    o->keys[2 * o->numKeys] = (uint32_t)key;
    o->keys[2 * o->numKeys + 1] = (uint32_t)m;
    o->numKeys++;
  }
  fclose(fp);
  /* sort (key, medium) pairs by key, then split into the lookup arrays */
  qsort(o->keys, (size_t)o->numKeys, 2 * sizeof(uint32_t), cmp_key);
  for (int q = 1; q < o->numKeys; q++)
    if (o->keys[2 * q] == o->keys[2 * q - 2]) vh_fail(s, "voxel file %s: key %u listed twice", file, o->keys[2 * q]);
  o->keyMedium = vh_alloc(s, (size_t)(o->numKeys ? o->numKeys : 1), sizeof(int));
  for (int q = 0; q < o->numKeys; q++) {
    o->keyMedium[q] = (int)o->keys[2 * q + 1];
    o->keys[q] = o->keys[2 * q];
  }

  size_t count = (size_t)o->n[0] * (size_t)o->n[1] * (size_t)o->n[2], want = count * (size_t)o->bytes;
  FILE *dp = vh_open_input(s, dataName, "rb");
  if (!dp) vh_fail(s, "cannot open voxel data file %s", dataName);
  o->data = malloc(want ? want : 1);
  if (!o->data) {
    fclose(dp);
    vh_fail(s, "out of host memory (%zu bytes of voxels)", want);
  }
  size_t got = fread(o->data, 1, want, dp);
  bool extra = fgetc(dp) != EOF;
  fclose(dp);
  if (got != want || extra)
    vh_fail(s, "voxel data file %s: expected exactly %zu bytes (%d x %d x %d keys of %d bytes)", dataName, want,
            o->n[0], o->n[1], o->n[2], o->bytes);
}

/* Record MB <node> <object> [theta [phi]]; returns the placeholder medium
 * index the block carries until vh_expand_voxels replaces it. */
int vh_place_voxel(vh_solve *s, int object, const int node[3], double theta, double phi)
{
  vh_grow(s, (void **)&s->voxelPlaces, &s->capVoxelPlaces, s->numVoxelPlaces + 1, sizeof(vh_voxel_place));
  vh_voxel_place *p = &s->voxelPlaces[s->numVoxelPlaces];
  p->object = object;
  for (int a = 0; a < 3; a++) p->node[a] = node[a];
  p->theta = theta;
  p->phi = phi;
  return -2 - s->numVoxelPlaces++;
}

/* sin and cos of an angle in degrees, exact at multiples of 90. */
static void sincos_deg(double deg, double *sn, double *cs)
{
  double q = deg / 90.0;
  if (q == floor(q) && fabs(q) < 1e9) {
    static const double S[4] = {0.0, 1.0, 0.0, -1.0}, C[4] = {1.0, 0.0, -1.0, 0.0};
    long r = ((long)q % 4 + 4) % 4;
    *sn = S[r];
    *cs = C[r];
    return;
  }
  double a = deg * (3.14159265358979323846 / 180.0);
  *sn = sin(a);
  *cs = cos(a);
}

typedef struct {
  double ct, st, cp, sp;
} rot;

/* mesh offset d (from the placement node) -> object coordinates */
static void to_object(const rot *r, const double d[3], double v[3])
{
  double wx = r->cp * d[0] + r->sp * d[1], wy = -r->sp * d[0] + r->cp * d[1], wz = d[2];
  v[0] = r->ct * wx - r->st * wz;
  v[1] = wy;
  v[2] = r->st * wx + r->ct * wz;
}

/* object coordinates -> mesh offset */
static void to_mesh(const rot *r, const double v[3], double d[3])
{
  double wx = r->ct * v[0] + r->st * v[2], wz = -r->st * v[0] + r->ct * v[2];
  d[0] = r->cp * wx - r->sp * v[1];
  d[1] = r->sp * wx + r->cp * v[1];
  d[2] = wz;
}

static const real *mesh_lines(const vh_solve *s, int a)
{
  return a == 0 ? s->xlines : a == 1 ? s->ylines : s->zlines;
}

/* Medium of every cell of one (i, j) column in [k0, k1], or -1. */
static void column_media(const vh_solve *s, const vh_voxel *o, const rot *r, const double P0[3], double cx,
                         double cy, int k0, int k1, int *med)
{
  const real *zl = mesh_lines(s, 2);
  for (int k = k0; k <= k1; k++) {
    double d[3] = {cx - P0[0], cy - P0[1], 0.5 * ((double)zl[k] + (double)zl[k + 1]) - P0[2]}, v[3];
    to_object(r, d, v);
    long idx[3];
    bool in = true;
    for (int a = 0; a < 3 && in; a++) {
      double t = floor((v[a] + o->ref[a]) / o->d[a]);
      in = t >= 0.0 && t < (double)o->n[a];
      idx[a] = in ? (long)t : 0;
    }
    med[k - k0] = in ? key_medium(o, voxel_key(o, ((size_t)idx[0] * (size_t)o->n[1] + (size_t)idx[1]) *
                                                      (size_t)o->n[2] + (size_t)idx[2]))
                     : -1;
  }
}

/* Cells [lo, hi] of axis a whose centres lie in [xmin, xmax]; false if none. */
static bool cell_range(const vh_solve *s, int a, double xmin, double xmax, int *lo, int *hi)
{
  const real *L = mesh_lines(s, a);
  *lo = -1;
  *hi = -2;
  for (int i = 0; i < s->numLines[a] - 1; i++) {
    double c = 0.5 * ((double)L[i] + (double)L[i + 1]);
    if (c < xmin || c > xmax) continue;
    if (*lo < 0) *lo = i;
    *hi = i;
  }
  return *lo >= 0;
}

static void push_block(vh_solve *s, vh_block **out, int *n, int *cap, const int b[6], int medium, FaceMask mask)
{
  vh_grow(s, (void **)out, cap, *n + 1, sizeof(vh_block));
  vh_block *blk = &(*out)[(*n)++];
  memcpy(blk->mbbox, b, 6 * sizeof(int));
  blk->medium = medium;
  blk->mask = mask;
}

static void expand_place(vh_solve *s, const vh_voxel_place *p, vh_block **out, int *n, int *cap)
{
  const vh_voxel *o = &s->voxels[p->object];
  rot r;
  sincos_deg(p->theta, &r.st, &r.ct);
  sincos_deg(p->phi, &r.sp, &r.cp);
  double P0[3];
  for (int a = 0; a < 3; a++) P0[a] = (double)mesh_lines(s, a)[p->node[a]];
  /* box around the placed object: its eight corners in the mesh */
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int c = 0; c < 8; c++) {
    double v[3], d[3];
    for (int a = 0; a < 3; a++) v[a] = ((c >> a) & 1 ? o->n[a] * o->d[a] : 0.0) - o->ref[a];
    to_mesh(&r, v, d);
    for (int a = 0; a < 3; a++) {
      lo[a] = fmin(lo[a], P0[a] + d[a]);
      hi[a] = fmax(hi[a], P0[a] + d[a]);
    }
  }
  int c0[3], c1[3];
  for (int a = 0; a < 3; a++)
    if (!cell_range(s, a, lo[a], hi[a], &c0[a], &c1[a])) return;
  int nk = c1[2] - c0[2] + 1;
  int *med = vh_alloc(s, (size_t)nk, sizeof(int));
  const real *xl = mesh_lines(s, 0), *yl = mesh_lines(s, 1);
  for (int i = c0[0]; i <= c1[0]; i++)
    for (int j = c0[1]; j <= c1[1]; j++) {
      column_media(s, o, &r, P0, 0.5 * ((double)xl[i] + (double)xl[i + 1]),
                   0.5 * ((double)yl[j] + (double)yl[j + 1]), c0[2], c1[2], med);
      for (int k = 0; k < nk;) {
        if (med[k] < 0) { k++; continue; }
        int e = k;
        while (e + 1 < nk && med[e + 1] == med[k]) e++;
        int b[6] = {i, i + 1, j, j + 1, c0[2] + k, c0[2] + e + 1};
        FaceMask mask = VOXEL_MASK_LOW;
        if (e + 1 >= nk || med[e + 1] < 0) mask |= VOXEL_MASK_ZHI;
        push_block(s, out, n, cap, b, med[k], mask);
        k = e + 1;
      }
    }
  free(med);
}

/* Replace every placement's placeholder block by its blocks.  Runs once the
 * mesh lines are known (the MB cards come before the line cards). */
void vh_expand_voxels(vh_solve *s)
{
  if (s->numVoxelPlaces == 0) return;
  vh_block *out = NULL;
  int n = 0, cap = 0;
  for (int b = 0; b < s->numBlocks; b++) {
    vh_block blk = s->blocks[b];
    int pl = -2 - blk.medium;
    if (blk.medium >= 0)
      push_block(s, &out, &n, &cap, blk.mbbox, blk.medium, blk.mask);
    else if (pl >= 0 && pl < s->numVoxelPlaces)
      expand_place(s, &s->voxelPlaces[pl], &out, &n, &cap);
    else {
      free(out);
      vh_fail(s, "internal: block %d has no medium", b);
    }
  }
  free(s->blocks);
  s->blocks = out;
  s->numBlocks = n;
  s->capBlocks = cap;
  s->numVoxelPlaces = 0;
}

void vh_free_voxels(vh_solve *s)
{
  for (int v = 0; v < s->numVoxels; v++) {
    free(s->voxels[v].data);
    free(s->voxels[v].keys);
    free(s->voxels[v].keyMedium);
  }
  free(s->voxels);
  free(s->voxelPlaces);
  s->voxels = NULL;
  s->voxelPlaces = NULL;
  s->numVoxels = s->numVoxelPlaces = 0;
}
