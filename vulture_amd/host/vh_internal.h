/* Internal helpers shared by the host set-up translation units. */
#ifndef VH_INTERNAL_H
#define VH_INTERNAL_H

#include <setjmp.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "vh.h"

/* Scan format for a `real` (the reference scans PML/SIBC numbers with %e into
 * real*, boundary.c:173,184; a double build needs %le). */
#if defined(VH_REAL_IS_FLOAT)
#define VH_RFMT "%e"
#else
#define VH_RFMT "%le"
#endif

/* Record an error and unwind to the active vh_read_mesh / vh_init call. */
void vh_fail(vh_solve *s, const char *fmt, ...) __attribute__((noreturn, format(printf, 2, 3)));

void *vh_alloc(vh_solve *s, size_t n, size_t size);
void vh_grow(vh_solve *s, void **arr, int *cap, int need, size_t elem);

/* Open an input file named in the mesh: relative names are tried in the
 * current directory first, then next to the mesh file. */
FILE *vh_open_input(vh_solve *s, const char *name, const char *mode);
FILE *vh_open_output(vh_solve *s, const char *name, const char *mode);

/* Card handlers (mesh.c:58-191 dispatch table). */
bool vh_parse_card(vh_solve *s, int card, char *line, FILE *fp);

/* Lookups by name. */
int vh_find_medium(const vh_solve *s, const char *name);
int vh_find_boundary(const vh_solve *s, const char *name);
int vh_find_waveform(const vh_solve *s, const char *name);
int vh_find_wire(const vh_solve *s, const char *name);

int vh_add_medium(vh_solve *s, const char *name, int type, real eps_r, real sigma, real mu_r,
                  int numPoles, const double *residues, const double *poles, const char *fileName);
int vh_add_boundary(vh_solve *s, const char *name, int type, int numLayers, int order, real n_eff,
                    real refCoeff, real kmax, const char *fileName, real S_TM[2][2], real S_TE[2][2]);
int vh_add_observer(vh_solve *s, const int mbbox[6], const int step[3], const char *name, int format,
                    int domain, int quantity, bool isInternal, int waveform);

void vh_set_medium_on_grid(vh_solve *s, const int gbbox[6], int medium, FaceMask mask);
void vh_calc_coeff(vh_solve *s, real *alpha, real *beta, real *gamma, double complex *dalpha,
                   double complex *dbeta, real dt, real eps_r, real sigma, real mu_r, int numPoles,
                   const double complex *residues, const double complex *poles);
void vh_face_mask_to_bool(bool includeBoundary[6], FaceMask mask);
void vh_face_of_bbox(int faceBox[6], const int bbox[6], int face);
void vh_offset_bbox(int out[6], const int bbox[6], const int offset[6]);
int vh_bbox_direction(const int bbox[6]);
bool vh_field_is_in_boundary(int field, int face);
real vh_numerical_phase_velocity(vh_solve *s, real theta, real phi);

/* Voxel objects (vh_voxel.c).  A placement's block carries the placeholder
 * medium vh_place_voxel returns until vh_expand_voxels replaces it. */
int vh_find_voxel(const vh_solve *s, const char *name);
void vh_load_voxel(vh_solve *s, const char *name, const char *file);
int vh_place_voxel(vh_solve *s, int object, const int node[3], double theta, double phi);
void vh_expand_voxels(vh_solve *s);
void vh_free_voxels(vh_solve *s);

/* SIBC filter helpers (filter.c:309-368). */
void vh_rc_from_pole_residue(vh_rcfilter *rc, int numPoles, double asymp, const double complex *poles,
                             const double complex *residues, double T);

#endif
