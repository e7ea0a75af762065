/*
 * lssph_oracle.c -- CPU oracle of the LS-SPH hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of
 * the product path.  It restates, loop by loop, the reference Fortran
 * (shobuzako-kensuke/LS-SPH-Benchmarks); every function cites the file:line
 * range it follows.  Differences from the reference are limited to
 *   - DGESV: the reference links LAPACK/MKL DGESV (unpinned version); here
 *     a textbook LU with partial pivoting (Golub & Van Loan, Alg. 3.4.1)
 *     with DGESV's interface semantics (B untouched when a pivot is 0);
 *   - the periodic box (params.periodic = 1), diff_coef/diff_source and the
 *     escaped-particle check, which the reference does not have;
 *   - cell lists are rebuilt with a deterministic counting sort whose
 *     within-cell order equals the reference's single-thread order
 *     (fluid in ascending index, then walls in ascending index);
 *   - SPH_MODEL 5 shifting uses the q = 14 basis: the reference's writes
 *     a_vec(15:20) past the end of its q = 14 array
 *     (shift_interpolate_mod.f90:174, guard ">= 5" where ">= 6" is meant),
 *     which flang rejects at compile time.
 * Operators: LS type A for SPH_MODEL 3-6 (q = 5, 9, 14, 20) and classical
 * SPH for SPH_MODEL 1-2; shifting with q = 2 for SPH_MODEL 1-2.
 * Per-particle loops are OpenMP-parallel; each iteration writes only its
 * own particle, so results do not depend on the thread count.
 */
#include "lssph_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592653589793 /* acos(-1.0d0) */
#define QMAX 20

struct orc_sim {
  lssph_params_t p;
  int n_int, n_ext, n_total, has_tem;
  int32_t *ptype;
  double *x, *y, *u, *v, *rho, *pre, *tem;
  double *x_old, *y_old, *u_old, *v_old, *rho_old, *tem_old;
  double *x_new, *y_new, *u_new, *v_new, *rho_new, *tem_new;
  double *rhs_u, *rhs_v, *rhs_rho, *rhs_tem;
  double *f_prev;
  /* virtual markers (VM_type, global_types.f90:108-114) */
  int32_t *vm_ptype;
  double *vm_x, *vm_y, *vm_u, *vm_v, *vm_pre, *vm_tem, *vm_nx, *vm_ny;
  /* cells (Cell_type, global_types.f90:119-124), 0-based cell ids */
  int nx, ny, n_cells, cells_ready;
  double inv_cw_x, inv_cw_y; /* periodic: cells per unit length */
  int32_t *cell_of;          /* idx_SP - 1 */
  int32_t *wall_cell;        /* fixed cells of the walls (init_cell) */
  int32_t *vm_cell;          /* idx_VM - 1 */
  int32_t *cell_start;       /* n_cells + 1 */
  int32_t *cell_list;        /* n_total */
  int32_t *cell_cursor;
  int32_t *nbr;              /* 9 per cell, -1 = none */
  double *nbr_sx, *nbr_sy;   /* periodic image shift per neighbor entry */
  char err[256];
};

static void set_err(orc_sim *s, const char *m) {
  snprintf(s->err, sizeof s->err, "%s", m);
}

void orc_set_threads(int32_t n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* Kernel functions: kernel_functions_mod.f90:26-81 (W), :87-146 (dW/dr).   */
/* ------------------------------------------------------------------------ */
static inline double ipow(double t, int n) {
  double r = 1.0;
  for (int i = 0; i < n; ++i) r *= t;
  return r;
}

double orc_kernel_w(int32_t kt, double r, double h) {
  double q = r / h, w;
  switch (kt) {
    case 1: /* cubic spline, :33-41 */
      if (0.0 <= q && q <= 1.0)
        w = ipow(2.0 - q, 3) - 4.0 * ipow(1.0 - q, 3);
      else if (1.0 < q && q <= 2.0)
        w = ipow(2.0 - q, 3);
      else
        w = 0.0;
      return 5.0 / (14.0 * ORC_PI * (h * h)) * w;
    case 2: /* quintic spline, :43-53 */
      if (0.0 <= q && q <= 1.0)
        w = ipow(3.0 - q, 5) - 6.0 * ipow(2.0 - q, 5) + 15.0 * ipow(1.0 - q, 5);
      else if (1.0 < q && q <= 2.0)
        w = ipow(3.0 - q, 5) - 6.0 * ipow(2.0 - q, 5);
      else if (2.0 < q && q <= 3.0)
        w = ipow(3.0 - q, 5);
      else
        w = 0.0;
      return 7.0 / (478.0 * ORC_PI * (h * h)) * w;
    case 3: /* Wendland C2, :55-61 */
      w = (0.0 <= q && q <= 2.0) ? ipow(1.0 - 0.5 * q, 4) * (1.0 + 2.0 * q) : 0.0;
      return 7.0 / (4.0 * ORC_PI * (h * h)) * w;
    case 4: /* Wendland C4, :63-69 */
      w = (0.0 <= q && q <= 2.0)
              ? ipow(1.0 - 0.5 * q, 6) * (1.0 + 3.0 * q + 35.0 * (q * q) / 12.0)
              : 0.0;
      return 9.0 / (4.0 * ORC_PI * (h * h)) * w;
    default: /* Wendland C6, :71-78 */
      w = (0.0 <= q && q <= 2.0)
              ? ipow(1.0 - 0.5 * q, 8) *
                    (1.0 + 4.0 * q + 25.0 * (q * q) / 4.0 + 4.0 * (q * q * q))
              : 0.0;
      return 39.0 / (14.0 * ORC_PI * (h * h)) * w;
  }
}

double orc_kernel_dw(int32_t kt, double r, double h) {
  double q = r / h, d;
  switch (kt) {
    case 1: /* :94-102 */
      if (0.0 <= q && q <= 1.0)
        d = ipow(2.0 - q, 2) - 4.0 * ipow(1.0 - q, 2);
      else if (1.0 < q && q <= 2.0)
        d = ipow(2.0 - q, 2);
      else
        d = 0.0;
      return 5.0 / (14.0 * ORC_PI * (h * h)) * d * (-3.0 / h);
    case 2: /* :104-114 */
      if (0.0 <= q && q <= 1.0)
        d = ipow(3.0 - q, 4) - 6.0 * ipow(2.0 - q, 4) + 15.0 * ipow(1.0 - q, 4);
      else if (1.0 < q && q <= 2.0)
        d = ipow(3.0 - q, 4) - 6.0 * ipow(2.0 - q, 4);
      else if (2.0 < q && q <= 3.0)
        d = ipow(3.0 - q, 4);
      else
        d = 0.0;
      return 7.0 / (478.0 * ORC_PI * (h * h)) * d * (-5.0 / h);
    case 3: /* :116-123 */
      d = (0.0 <= q && q <= 2.0)
              ? -2.0 * ipow(1.0 - 0.5 * q, 3) * (1.0 + 2.0 * q) +
                    2.0 * ipow(1.0 - 0.5 * q, 4)
              : 0.0;
      return 7.0 / (4.0 * ORC_PI * (h * h)) * d / h;
    case 4: /* :125-132 */
      d = (0.0 <= q && q <= 2.0)
              ? -3.0 * ipow(1.0 - 0.5 * q, 5) *
                        (1.0 + 3.0 * q + 35.0 * (q * q) / 12.0) +
                    1.0 * ipow(1.0 - 0.5 * q, 6) * (3.0 + 35.0 * q / 6.0)
              : 0.0;
      return 9.0 / (4.0 * ORC_PI * (h * h)) * d / h;
    default: /* :134-143 */
      d = (0.0 <= q && q <= 2.0)
              ? -4.0 * ipow(1.0 - 0.5 * q, 7) *
                        (1.0 + 4.0 * q + 25.0 * (q * q) / 4.0 + 4.0 * (q * q * q)) +
                    1.0 * ipow(1.0 - 0.5 * q, 8) *
                        (4.0 + 25.0 * q / 2.0 + 12.0 * (q * q))
              : 0.0;
      return 39.0 / (14.0 * ORC_PI * (h * h)) * d / h;
  }
}

/* ------------------------------------------------------------------------ */
/* DGESV(N, NRHS, A, LDA, IPIV, B, LDB, INFO) as called at LSSPH_A_mod.f90: */
/* 211, shift_interpolate_mod.f90:358, interpolate_LSSPH_B_mod.f90:237.     */
/* LU with partial pivoting (first index of the largest |a| in the column), */
/* full-row interchanges, then forward/back substitution.  INFO = j > 0 if  */
/* U(j,j) is exactly zero; B is then left as it was (LAPACK dgesv).         */
/* ------------------------------------------------------------------------ */
int32_t orc_dgesv(int32_t n, int32_t nrhs, double *a, int32_t *ipiv,
                  double *b) {
  int32_t info = 0;
  for (int j = 0; j < n; ++j) {
    int p = j;
    double amax = fabs(a[j + j * n]);
    for (int i = j + 1; i < n; ++i) {
      double t = fabs(a[i + j * n]);
      if (t > amax) {
        amax = t;
        p = i;
      }
    }
    ipiv[j] = p;
    if (a[p + j * n] != 0.0) {
      if (p != j)
        for (int c = 0; c < n; ++c) {
          double t = a[j + c * n];
          a[j + c * n] = a[p + c * n];
          a[p + c * n] = t;
        }
      for (int i = j + 1; i < n; ++i) a[i + j * n] /= a[j + j * n];
    } else if (info == 0) {
      info = j + 1;
    }
    for (int c = j + 1; c < n; ++c)
      for (int i = j + 1; i < n; ++i) a[i + c * n] -= a[i + j * n] * a[j + c * n];
  }
  if (info) return info;
  for (int r = 0; r < nrhs; ++r) {
    double *bb = b + r * n;
    for (int j = 0; j < n; ++j)
      if (ipiv[j] != j) {
        double t = bb[j];
        bb[j] = bb[ipiv[j]];
        bb[ipiv[j]] = t;
      }
    for (int j = 0; j < n; ++j)
      for (int i = j + 1; i < n; ++i) bb[i] -= bb[j] * a[i + j * n];
    for (int j = n - 1; j >= 0; --j) {
      bb[j] /= a[j + j * n];
      for (int i = 0; i < j; ++i) bb[i] -= bb[j] * a[i + j * n];
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Taylor basis of LS type A (LSSPH_A_mod.f90:140-170 and                   */
/* shift_interpolate_mod.f90:266-297 / :372-403): q = 5, 9, 14, 20.         */
/* ------------------------------------------------------------------------ */
static inline void basis_a(int q, double x, double y, const double *ih, double *a) {
  a[0] = x * ih[0];
  a[1] = y * ih[0];
  if (q < 5) return;
  a[2] = 0.5 * (x * x) * ih[1];
  a[3] = 0.5 * (y * y) * ih[1];
  a[4] = x * y * ih[1];
  if (q < 9) return;
  a[5] = 1.0 / 6.0 * (x * x * x) * ih[2];
  a[6] = 1.0 / 6.0 * (y * y * y) * ih[2];
  a[7] = 0.5 * (x * x) * y * ih[2];
  a[8] = 0.5 * (y * y) * x * ih[2];
  if (q < 14) return;
  a[9] = 1.0 / 24.0 * (x * x * x * x) * ih[3];
  a[10] = 1.0 / 24.0 * (y * y * y * y) * ih[3];
  a[11] = 1.0 / 6.0 * (x * x * x) * y * ih[3];
  a[12] = 1.0 / 6.0 * (y * y * y) * x * ih[3];
  a[13] = 1.0 / 4.0 * (x * x) * (y * y) * ih[3];
  if (q < 20) return;
  a[14] = 1.0 / 120.0 * (x * x * x * x * x) * ih[4];
  a[15] = 1.0 / 120.0 * (y * y * y * y * y) * ih[4];
  a[16] = 1.0 / 24.0 * (x * x * x * x) * y * ih[4];
  a[17] = 1.0 / 24.0 * (y * y * y * y) * x * ih[4];
  a[18] = 1.0 / 12.0 * (x * x * x) * (y * y) * ih[4];
  a[19] = 1.0 / 12.0 * (y * y * y) * (x * x) * ih[4];
}

static int q_of_model(int sph_model, int *k) {
  switch (sph_model) {
    case 1:
    case 2: *k = 1; return 2; /* shifting only, shift_interpolate_mod.f90:43-44 */
    case 4: *k = 3; return 9;
    case 5: *k = 4; return 14;
    case 6: *k = 5; return 20;
    default: *k = 2; return 5;
  }
}

/* ------------------------------------------------------------------------ */
/* Creation / state                                                        */
/* ------------------------------------------------------------------------ */
#define ALLOC(ptr, n) \
  do { (ptr) = calloc((size_t)((n) > 0 ? (n) : 1), sizeof(*(ptr))); } while (0)

orc_sim *orc_create(const lssph_params_t *p, int32_t n_int, int32_t n_ext) {
  orc_sim *s = calloc(1, sizeof(orc_sim));
  if (!s) return NULL;
  s->p = *p;
  s->n_int = n_int;
  s->n_ext = n_ext;
  s->n_total = n_int + n_ext;
  s->has_tem = (p->target_problem == 4);
  int nt = s->n_total, ni = n_int, ne = n_ext;
  ALLOC(s->ptype, nt);
  ALLOC(s->x, nt); ALLOC(s->y, nt); ALLOC(s->u, nt); ALLOC(s->v, nt);
  ALLOC(s->rho, nt); ALLOC(s->pre, nt); ALLOC(s->tem, nt);
  ALLOC(s->x_old, ni); ALLOC(s->y_old, ni); ALLOC(s->u_old, ni);
  ALLOC(s->v_old, ni); ALLOC(s->rho_old, ni); ALLOC(s->tem_old, ni);
  ALLOC(s->x_new, ni); ALLOC(s->y_new, ni); ALLOC(s->u_new, ni);
  ALLOC(s->v_new, ni); ALLOC(s->rho_new, ni); ALLOC(s->tem_new, ni);
  ALLOC(s->rhs_u, ni); ALLOC(s->rhs_v, ni); ALLOC(s->rhs_rho, ni);
  ALLOC(s->rhs_tem, ni); ALLOC(s->f_prev, ni);
  for (int i = 0; i < ni; ++i) s->f_prev[i] = 1.0; /* setup_closed_box_mod.f90:180-183 */
  ALLOC(s->vm_ptype, ne); ALLOC(s->vm_x, ne); ALLOC(s->vm_y, ne);
  ALLOC(s->vm_u, ne); ALLOC(s->vm_v, ne); ALLOC(s->vm_pre, ne);
  ALLOC(s->vm_tem, ne); ALLOC(s->vm_nx, ne); ALLOC(s->vm_ny, ne);
  ALLOC(s->cell_of, nt); ALLOC(s->wall_cell, ne); ALLOC(s->vm_cell, ne);
  ALLOC(s->cell_list, nt);
  return s;
}

void orc_destroy(orc_sim *s) {
  if (!s) return;
  void *ptrs[] = {s->ptype, s->x, s->y, s->u, s->v, s->rho, s->pre, s->tem,
                  s->x_old, s->y_old, s->u_old, s->v_old, s->rho_old, s->tem_old,
                  s->x_new, s->y_new, s->u_new, s->v_new, s->rho_new, s->tem_new,
                  s->rhs_u, s->rhs_v, s->rhs_rho, s->rhs_tem, s->f_prev,
                  s->vm_ptype, s->vm_x, s->vm_y, s->vm_u, s->vm_v, s->vm_pre,
                  s->vm_tem, s->vm_nx, s->vm_ny, s->cell_of, s->wall_cell,
                  s->vm_cell, s->cell_start, s->cell_list, s->cell_cursor,
                  s->nbr, s->nbr_sx, s->nbr_sy};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
  free(s);
}

const char *orc_error(const orc_sim *s) { return s ? s->err : "null"; }

int32_t orc_set_state(orc_sim *s, const int32_t *ptype, const double *x,
                      const double *y, const double *u, const double *v,
                      const double *rho, const double *pre, const double *tem) {
  size_t nd = sizeof(double) * (size_t)s->n_total;
  memcpy(s->ptype, ptype, sizeof(int32_t) * (size_t)s->n_total);
  memcpy(s->x, x, nd); memcpy(s->y, y, nd); memcpy(s->u, u, nd);
  memcpy(s->v, v, nd); memcpy(s->rho, rho, nd); memcpy(s->pre, pre, nd);
  if (s->has_tem && tem) memcpy(s->tem, tem, nd);
  return 0;
}

int32_t orc_get_state(const orc_sim *s, int32_t *ptype, double *x, double *y,
                      double *u, double *v, double *rho, double *pre,
                      double *tem) {
  size_t nd = sizeof(double) * (size_t)s->n_total;
  if (ptype) memcpy(ptype, s->ptype, sizeof(int32_t) * (size_t)s->n_total);
  if (x) memcpy(x, s->x, nd);
  if (y) memcpy(y, s->y, nd);
  if (u) memcpy(u, s->u, nd);
  if (v) memcpy(v, s->v, nd);
  if (rho) memcpy(rho, s->rho, nd);
  if (pre) memcpy(pre, s->pre, nd);
  if (tem && s->has_tem) memcpy(tem, s->tem, nd);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* init_virtual_markers_mod.f90:61-137: VM k mirrors wall n_int + k.        */
/* ------------------------------------------------------------------------ */
int32_t orc_init_vm(orc_sim *s) {
  const lssph_params_t *p = &s->p;
  double x_min = 0.0 + p->wl_thick, x_max = p->len_x + p->wl_thick;
  double y_min = 0.0 + p->wl_thick, y_max = p->len_y + p->wl_thick;
  double r2 = 1.0 / sqrt(2.0);
  for (int k = 0; k < s->n_ext; ++k) {
    int w = s->n_int + k;
    int t = s->ptype[w];
    double wx = s->x[w], wy = s->y[w];
    s->vm_ptype[k] = t;
    switch (t) {
      case 1: s->vm_x[k] = wx; s->vm_y[k] = 2.0 * y_min - wy; s->vm_nx[k] = 0.0; s->vm_ny[k] = -1.0; break;
      case 2: s->vm_x[k] = wx; s->vm_y[k] = 2.0 * y_max - wy; s->vm_nx[k] = 0.0; s->vm_ny[k] = 1.0; break;
      case 3: s->vm_x[k] = 2.0 * x_min - wx; s->vm_y[k] = wy; s->vm_nx[k] = -1.0; s->vm_ny[k] = 0.0; break;
      case 4: s->vm_x[k] = 2.0 * x_max - wx; s->vm_y[k] = wy; s->vm_nx[k] = 1.0; s->vm_ny[k] = 0.0; break;
      case 5: s->vm_x[k] = 2.0 * x_min - wx; s->vm_y[k] = 2.0 * y_min - wy; s->vm_nx[k] = -r2; s->vm_ny[k] = -r2; break;
      case 6: s->vm_x[k] = 2.0 * x_max - wx; s->vm_y[k] = 2.0 * y_min - wy; s->vm_nx[k] = r2; s->vm_ny[k] = -r2; break;
      case 7: s->vm_x[k] = 2.0 * x_min - wx; s->vm_y[k] = 2.0 * y_max - wy; s->vm_nx[k] = -r2; s->vm_ny[k] = r2; break;
      case 8: s->vm_x[k] = 2.0 * x_max - wx; s->vm_y[k] = 2.0 * y_max - wy; s->vm_nx[k] = r2; s->vm_ny[k] = r2; break;
      default: set_err(s, "init_vm: wall particle with ptype outside 1..8"); return LSSPH_E_ARG;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Cell list: init_cell_mod.f90:36-182 and update_cell_mod.f90:40-93.       */
/* ------------------------------------------------------------------------ */
static int fluid_cell(const orc_sim *s, double x, double y) {
  if (s->p.periodic) {
    int cx = (int)floor(x * s->inv_cw_x), cy = (int)floor(y * s->inv_cw_y);
    if (cx == s->nx) cx = s->nx - 1;
    if (cy == s->ny) cy = s->ny - 1;
    if (cx < 0 || cx >= s->nx || cy < 0 || cy >= s->ny) return -1;
    return cx + cy * s->nx;
  }
  /* update_cell_mod.f90:53-55: ceiling(x * inv_h_eff) */
  double inv = 1.0 / s->p.h_eff;
  int cx = (int)ceil(x * inv), cy = (int)ceil(y * inv);
  if (cx < 1 || cx > s->nx || cy < 1 || cy > s->ny) return -1;
  return (cx - 1) + (cy - 1) * s->nx;
}

static int div_cell(const orc_sim *s, double x, double y) {
  /* init_cell_mod.f90:76-78 and :170-172: ceiling(x / h_eff) */
  int cx = (int)ceil(x / s->p.h_eff), cy = (int)ceil(y / s->p.h_eff);
  if (cx < 1 || cx > s->nx || cy < 1 || cy > s->ny) return -1;
  return (cx - 1) + (cy - 1) * s->nx;
}

int32_t orc_update_cell(orc_sim *s) {
  if (!s->cells_ready) {
    set_err(s, "update_cell before init_cell");
    return LSSPH_E_STATE;
  }
  int nc = s->n_cells;
  memset(s->cell_start, 0, sizeof(int32_t) * (size_t)(nc + 1));
  for (int i = 0; i < s->n_int; ++i) {
    int c = fluid_cell(s, s->x[i], s->y[i]);
    if (c < 0) {
      snprintf(s->err, sizeof s->err, "particle %d left the cell grid (x=%g, y=%g)", i, s->x[i], s->y[i]);
      return LSSPH_E_DOMAIN;
    }
    s->cell_of[i] = c;
    s->cell_start[c + 1]++;
  }
  for (int k = 0; k < s->n_ext; ++k) s->cell_start[s->wall_cell[k] + 1]++;
  for (int c = 0; c < nc; ++c) s->cell_start[c + 1] += s->cell_start[c];
  memcpy(s->cell_cursor, s->cell_start, sizeof(int32_t) * (size_t)nc);
  /* fluid in ascending index (single-thread atomic-capture order, :61-67),
     then the fixed wall list appended (:76-92) */
  for (int i = 0; i < s->n_int; ++i) s->cell_list[s->cell_cursor[s->cell_of[i]]++] = i;
  for (int k = 0; k < s->n_ext; ++k)
    s->cell_list[s->cell_cursor[s->wall_cell[k]]++] = s->n_int + k;
  return 0;
}

int32_t orc_init_cell(orc_sim *s) {
  const lssph_params_t *p = &s->p;
  if (p->periodic) {
    s->nx = (int)floor(p->len_x / p->h_eff);
    s->ny = (int)floor(p->len_y / p->h_eff);
    if (s->nx < 3 || s->ny < 3) {
      set_err(s, "periodic box needs at least 3 cells per direction");
      return LSSPH_E_ARG;
    }
    s->inv_cw_x = (double)s->nx / p->len_x;
    s->inv_cw_y = (double)s->ny / p->len_y;
  } else {
    /* init_cell_mod.f90:36-41 */
    double xm = p->len_x + 2.0 * p->wl_thick, ym = p->len_y + 2.0 * p->wl_thick;
    s->nx = (int)ceil(xm / p->h_eff);
    s->ny = (int)ceil(ym / p->h_eff);
  }
  int nc = s->n_cells = s->nx * s->ny;
  free(s->cell_start); free(s->cell_cursor); free(s->nbr); free(s->nbr_sx); free(s->nbr_sy);
  ALLOC(s->cell_start, nc + 1);
  ALLOC(s->cell_cursor, nc + 1);
  ALLOC(s->nbr, 9 * nc);
  ALLOC(s->nbr_sx, 9 * nc);
  ALLOC(s->nbr_sy, 9 * nc);
  /* 9 neighbor cells, order dy = -1..1 outer, dx = -1..1 inner (:118-154) */
  for (int c = 0; c < nc; ++c) {
    int cx = c % s->nx, cy = c / s->nx, k = 0;
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx, ++k) {
        int xx = cx + dx, yy = cy + dy;
        double sx = 0.0, sy = 0.0;
        if (p->periodic) {
          if (xx < 0) { xx += s->nx; sx = -p->len_x; }
          if (xx >= s->nx) { xx -= s->nx; sx = p->len_x; }
          if (yy < 0) { yy += s->ny; sy = -p->len_y; }
          if (yy >= s->ny) { yy -= s->ny; sy = p->len_y; }
          s->nbr[9 * c + k] = xx + yy * s->nx;
        } else {
          s->nbr[9 * c + k] =
              (xx < 0 || xx >= s->nx || yy < 0 || yy >= s->ny) ? -1 : xx + yy * s->nx;
        }
        s->nbr_sx[9 * c + k] = sx;
        s->nbr_sy[9 * c + k] = sy;
      }
  }
  /* wall cells (:70-86, :93-97) and VM cells (:166-176) use x / h_eff */
  for (int k = 0; k < s->n_ext; ++k) {
    int w = s->n_int + k;
    int c = div_cell(s, s->x[w], s->y[w]);
    int cv = div_cell(s, s->vm_x[k], s->vm_y[k]);
    if (c < 0 || cv < 0) {
      set_err(s, "wall particle or virtual marker outside the cell grid");
      return LSSPH_E_DOMAIN;
    }
    s->wall_cell[k] = c;
    s->cell_of[w] = c;
    s->vm_cell[k] = cv;
  }
  s->cells_ready = 1;
  return orc_update_cell(s);
}

int32_t orc_get_cell_index(const orc_sim *s, int32_t *idx_sp) {
  for (int i = 0; i < s->n_total; ++i) idx_sp[i] = s->cell_of[i] + 1;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* LSSPH_A (LSSPH_A_mod.f90:26-271) for particle me.                        */
/* ------------------------------------------------------------------------ */
static void lssph_a(orc_sim *s, int me) {
  const lssph_params_t *p = &s->p;
  int k, q = q_of_model(p->sph_model, &k);
  const int nb = s->has_tem ? 5 : 4; /* :58-62 */
  double ih[5], a[QMAX], M[QMAX * QMAX], b[QMAX * 5];
  int32_t ipiv[QMAX];
  memset(M, 0, sizeof M);
  memset(b, 0, sizeof b);
  double h = p->h, h_eff_sq = p->h_eff * p->h_eff, m = p->sp_mass;
  double my_x = s->x[me], my_y = s->y[me], my_u = s->u[me], my_v = s->v[me];
  double my_rho = s->rho[me], my_pre = s->pre[me], my_tem = s->has_tem ? s->tem[me] : 0.0;
  ih[0] = 1.0 / h;
  for (int i = 1; i < k; ++i) ih[i] = ih[i - 1] / h;
  int num = 0, c0 = s->cell_of[me];
  for (int nc = 0; nc < 9; ++nc) {
    int c = s->nbr[9 * c0 + nc];
    if (c < 0) continue;
    double sx = s->nbr_sx[9 * c0 + nc], sy = s->nbr_sy[9 * c0 + nc];
    for (int t = s->cell_start[c]; t < s->cell_start[c + 1]; ++t) {
      int j = s->cell_list[t];
      if (j == me) continue;
      double xj = (s->x[j] + sx) - my_x, yj = (s->y[j] + sy) - my_y;
      double r2 = xj * xj + yj * yj;
      if (r2 > h_eff_sq) continue;
      double w = orc_kernel_w(p->kernel_type, sqrt(r2), h);
      ++num;
      double vw = (m / s->rho[j]) * w;
      basis_a(q, xj, yj, ih, a);
      double df[5] = {s->u[j] - my_u, s->v[j] - my_v, s->rho[j] - my_rho,
                      s->pre[j] - my_pre, s->has_tem ? s->tem[j] - my_tem : 0.0};
      for (int i = 0; i < q; ++i) {
        double vwa = vw * a[i];
        for (int f = 0; f < nb; ++f) b[i + f * q] += vwa * df[f];
        for (int r = 0; r <= i; ++r) M[r + i * q] += vwa * a[r];
      }
    }
  }
  for (int i = 0; i < q - 1; ++i)
    for (int r = i + 1; r < q; ++r) M[r + i * q] = M[i + r * q];
  double d[5][5] = {{0}};
  if (num >= q) {
    orc_dgesv(q, nb, M, ipiv, b);
    for (int f = 0; f < nb; ++f) {
      d[f][0] = b[0 + f * q] * ih[0];
      d[f][1] = b[1 + f * q] * ih[0];
      d[f][2] = b[2 + f * q] * ih[1];
      d[f][3] = b[3 + f * q] * ih[1];
      d[f][4] = b[4 + f * q] * ih[1];
    }
  }
  /* RHS, :246-268 */
  if (p->target_problem == 1) {
    s->rhs_u[me] = p->diff_coef * (d[0][2] + d[0][3]) +
                   p->diff_source * (2.0 * ORC_PI * ORC_PI * sin(ORC_PI * (my_x - p->wl_thick)) *
                                     cos(ORC_PI * (my_y - p->wl_thick)));
    return;
  }
  double delta = p->coe_delta_sph * p->sound_speed * h * p->zeta_rsst;
  double r_rho = -my_rho * (d[0][0] + d[1][1]) + delta * (d[2][2] + d[2][3]);
  double r_u = -d[3][0] / my_rho + (p->vis_ref / my_rho) * (d[0][2] + d[0][3]);
  double r_v = -d[3][1] / my_rho + (p->vis_ref / my_rho) * (d[1][2] + d[1][3]);
  r_rho = r_rho / (p->zeta_rsst * p->zeta_rsst);
  if (p->target_problem == 4) {
    r_v = r_v + p->alpha_ref * (my_tem - p->tem_ave) * p->gravity;
    r_u = r_u / (p->xi_vim * p->xi_vim);
    r_v = r_v / (p->xi_vim * p->xi_vim);
    s->rhs_tem[me] = p->thermal_dif * (d[4][2] + d[4][3]);
  }
  s->rhs_rho[me] = r_rho;
  s->rhs_u[me] = r_u;
  s->rhs_v[me] = r_v;
}

/* ------------------------------------------------------------------------ */
/* classical_SPH (classical_SPH_mod.f90:26-170) for particle me: kernel      */
/* gradient sums with x_ij = x_i - x_j; SPH_MODEL 1 sums the pressures in    */
/* the gradient (:115-117), SPH_MODEL 2 differences them (:118-120).        */
/* ------------------------------------------------------------------------ */
static void classical_sph(orc_sim *s, int me) {
  const lssph_params_t *p = &s->p;
  const int sum_model = (p->sph_model == 1);
  double h = p->h, h_eff_sq = p->h_eff * p->h_eff, m = p->sp_mass;
  double my_x = s->x[me], my_y = s->y[me], my_u = s->u[me], my_v = s->v[me];
  double my_rho = s->rho[me], my_pre = s->pre[me], my_tem = s->has_tem ? s->tem[me] : 0.0;
  double du_dx = 0.0, dv_dy = 0.0, dp_dx = 0.0, dp_dy = 0.0;
  double lap_u = 0.0, lap_v = 0.0, lap_tem = 0.0, lap_rho = 0.0;
  int c0 = s->cell_of[me];
  for (int nc = 0; nc < 9; ++nc) {
    int c = s->nbr[9 * c0 + nc];
    if (c < 0) continue;
    double sx = s->nbr_sx[9 * c0 + nc], sy = s->nbr_sy[9 * c0 + nc];
    for (int t = s->cell_start[c]; t < s->cell_start[c + 1]; ++t) {
      int j = s->cell_list[t];
      if (j == me) continue;
      double x_ij = my_x - (s->x[j] + sx), y_ij = my_y - (s->y[j] + sy);
      double r2 = x_ij * x_ij + y_ij * y_ij;
      if (r2 > h_eff_sq) continue;
      double r = sqrt(r2);
      double dwdr = orc_kernel_dw(p->kernel_type, r, h);
      double vj = m / s->rho[j];
      double dwx = dwdr * x_ij / r * vj;
      double dwy = dwdr * y_ij / r * vj;
      du_dx = du_dx + dwx * (s->u[j] - my_u);
      dv_dy = dv_dy + dwy * (s->v[j] - my_v);
      double pp = sum_model ? s->pre[j] + my_pre : s->pre[j] - my_pre;
      dp_dx = dp_dx + dwx * pp;
      dp_dy = dp_dy + dwy * pp;
      double g = 2.0 * vj * dwdr / r;
      lap_u = lap_u + g * (my_u - s->u[j]);
      lap_v = lap_v + g * (my_v - s->v[j]);
      if (s->has_tem) lap_tem = lap_tem + g * (my_tem - s->tem[j]);
      lap_rho = lap_rho + g * (my_rho - s->rho[j]);
    }
  }
  /* :139-168 */
  if (p->target_problem == 1) {
    s->rhs_u[me] = p->diff_coef * lap_u +
                   p->diff_source * (2.0 * ORC_PI * ORC_PI * sin(ORC_PI * (my_x - p->wl_thick)) *
                                     cos(ORC_PI * (my_y - p->wl_thick)));
    return;
  }
  double delta = p->coe_delta_sph * p->sound_speed * h * p->zeta_rsst;
  double r_rho = -my_rho * (du_dx + dv_dy) + delta * lap_rho;
  double r_u = -dp_dx / my_rho + (p->vis_ref / my_rho) * lap_u;
  double r_v = -dp_dy / my_rho + (p->vis_ref / my_rho) * lap_v;
  r_rho = r_rho / (p->zeta_rsst * p->zeta_rsst);
  if (p->target_problem == 4) {
    r_v = r_v + p->alpha_ref * (my_tem - p->tem_ave) * p->gravity;
    r_u = r_u / (p->xi_vim * p->xi_vim);
    r_v = r_v / (p->xi_vim * p->xi_vim);
    s->rhs_tem[me] = p->thermal_dif * lap_tem;
  }
  s->rhs_rho[me] = r_rho;
  s->rhs_u[me] = r_u;
  s->rhs_v[me] = r_v;
}

/* calculate_RHS_mod.f90:36-50: classical SPH for SPH_MODEL <= 2, LS type A
   otherwise. */
int32_t orc_calculate_rhs(orc_sim *s) {
  if (!s->cells_ready) { set_err(s, "calculate_rhs before init_cell"); return LSSPH_E_STATE; }
  const int classical = s->p.sph_model <= 2;
#pragma omp parallel for schedule(static)
  for (int me = 0; me < s->n_int; ++me) {
    if (classical)
      classical_sph(s, me);
    else
      lssph_a(s, me);
  }
  return 0;
}

int32_t orc_get_rhs(const orc_sim *s, double *ru, double *rv, double *rrho, double *rtem) {
  size_t nd = sizeof(double) * (size_t)s->n_int;
  if (ru) memcpy(ru, s->rhs_u, nd);
  if (rv) memcpy(rv, s->rhs_v, nd);
  if (rrho) memcpy(rrho, s->rhs_rho, nd);
  if (rtem && s->has_tem) memcpy(rtem, s->rhs_tem, nd);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Ghost walls: interpolate_LSSPH_B_mod.f90:24-178, interpolate_CSPH_mod.f90 */
/* :24-130, extrapolate_VM_to_WL_mod.f90:23-275,                            */
/* extrapolate_diff_test_mod.f90:24-106, ghost_mod.f90:51-76.               */
/* ------------------------------------------------------------------------ */
static void interpolate_b(orc_sim *s, int me) {
  const lssph_params_t *p = &s->p;
  int q = (p->wall_model == 3) ? 6 : 3;
  int nb = s->has_tem ? 4 : 3;
  double a[6] = {1.0, 0, 0, 0, 0, 0}, M[36], b[24];
  int32_t ipiv[6];
  memset(M, 0, sizeof M);
  memset(b, 0, sizeof b);
  double h = p->h, h_eff_sq = p->h_eff * p->h_eff, m = p->sp_mass;
  double my_x = s->vm_x[me], my_y = s->vm_y[me];
  double inv_h = 1.0 / h, inv_h2 = inv_h * inv_h;
  int num = 0, c0 = s->vm_cell[me];
  for (int nc = 0; nc < 9; ++nc) {
    int c = s->nbr[9 * c0 + nc];
    if (c < 0) continue;
    for (int t = s->cell_start[c]; t < s->cell_start[c + 1]; ++t) {
      int j = s->cell_list[t];
      if (s->ptype[j] != 0) continue;
      double xj = s->x[j] - my_x, yj = s->y[j] - my_y;
      double r2 = xj * xj + yj * yj;
      if (r2 > h_eff_sq) continue;
      double w = orc_kernel_w(p->kernel_type, sqrt(r2), h);
      ++num;
      double vw = (m / s->rho[j]) * w;
      a[1] = xj * inv_h;
      a[2] = yj * inv_h;
      if (q == 6) {
        a[3] = xj * xj * 0.5 * inv_h2;
        a[4] = xj * yj * inv_h2;
        a[5] = yj * yj * 0.5 * inv_h2;
      }
      double fv[4] = {s->u[j], s->v[j], s->pre[j], s->has_tem ? s->tem[j] : 0.0};
      for (int i = 0; i < q; ++i) {
        double vwa = vw * a[i];
        for (int f = 0; f < nb; ++f) b[i + f * q] += vwa * fv[f];
        for (int r = 0; r <= i; ++r) M[r + i * q] += vwa * a[r];
      }
    }
  }
  for (int i = 0; i < q - 1; ++i)
    for (int r = i + 1; r < q; ++r) M[r + i * q] = M[i + r * q];
  if (num >= q) {
    orc_dgesv(q, nb, M, ipiv, b);
    s->vm_u[me] = b[0];
    s->vm_v[me] = b[q];
    s->vm_pre[me] = b[2 * q];
    if (s->has_tem) s->vm_tem[me] = b[3 * q];
  } else {
    s->vm_u[me] = s->vm_v[me] = s->vm_pre[me] = 0.0;
    if (s->has_tem) s->vm_tem[me] = 0.0;
  }
}

static void interpolate_csph(orc_sim *s, int me) {
  const lssph_params_t *p = &s->p;
  double h = p->h, h_eff_sq = p->h_eff * p->h_eff, m = p->sp_mass;
  double my_x = s->vm_x[me], my_y = s->vm_y[me];
  double su = 0, sv = 0, sp = 0, st = 0, sw = 0;
  int num = 0, c0 = s->vm_cell[me];
  for (int nc = 0; nc < 9; ++nc) {
    int c = s->nbr[9 * c0 + nc];
    if (c < 0) continue;
    for (int t = s->cell_start[c]; t < s->cell_start[c + 1]; ++t) {
      int j = s->cell_list[t];
      if (s->ptype[j] != 0) continue;
      double xij = my_x - s->x[j], yij = my_y - s->y[j];
      double r2 = xij * xij + yij * yij;
      if (r2 > h_eff_sq) continue;
      double w = orc_kernel_w(p->kernel_type, sqrt(r2), h);
      ++num;
      double vw = (m / s->rho[j]) * w;
      sw = sw + vw;
      su = su + vw * s->u[j];
      sv = sv + vw * s->v[j];
      sp = sp + vw * s->pre[j];
      if (s->has_tem) st = st + vw * s->tem[j];
    }
  }
  if (num > 0 && sw > 1.0e-12) {
    s->vm_u[me] = su / sw;
    s->vm_v[me] = sv / sw;
    s->vm_pre[me] = sp / sw;
    if (s->has_tem) s->vm_tem[me] = st / sw;
  } else {
    s->vm_u[me] = s->vm_v[me] = s->vm_pre[me] = 0.0;
    if (s->has_tem) s->vm_tem[me] = 0.0;
  }
}

/* Pressure of an isothermal wall side with the hydrostatic correction
   (extrapolate_VM_to_WL_mod.f90:57-61,105-109,209-213,250-254). */
static double wall_pre(const orc_sim *s, int me, int you, int iso, double tw) {
  const lssph_params_t *p = &s->p;
  if (s->has_tem && iso == 1)
    return s->vm_pre[me] + p->rho_ref * p->alpha_ref * (tw - p->tem_ave) * p->gravity *
                               (s->y[you] - s->vm_y[me]);
  return s->vm_pre[me];
}

static double wall_tem(const orc_sim *s, int me, int bc, double tw) {
  return (bc == 1) ? 2.0 * tw - s->vm_tem[me] : s->vm_tem[me];
}

static void extrapolate_vm_to_wl(orc_sim *s, int me) {
  const lssph_params_t *p = &s->p;
  int you = s->n_int + me;
  double nx = s->vm_nx[me], ny = s->vm_ny[me];
  double vu = s->vm_u[me], vv = s->vm_v[me];
  int t = s->vm_ptype[me], bc;
  if (t >= 1 && t <= 4) {
    bc = (t == 1) ? p->u_bc_bottom : (t == 2) ? p->u_bc_top : (t == 3) ? p->u_bc_left : p->u_bc_right;
    if (bc == 1) { /* free-slip */
      double un = vu * nx + vv * ny;
      s->u[you] = vu - 2.0 * un * nx;
      s->v[you] = vv - 2.0 * un * ny;
    } else if (bc == 2) { /* no-slip */
      s->u[you] = -vu;
      s->v[you] = -vv;
    } else if (bc == 3 && t == 2) { /* moving lid (top only) */
      s->u[you] = 2.0 * p->u_top - vu;
      s->v[you] = -vv;
    }
  }
  switch (t) {
    case 1:
      s->pre[you] = wall_pre(s, me, you, p->tem_bc_bottom, p->tem_bottom);
      if (s->has_tem) s->tem[you] = wall_tem(s, me, p->tem_bc_bottom, p->tem_bottom);
      break;
    case 2:
      s->pre[you] = wall_pre(s, me, you, p->tem_bc_top, p->tem_top);
      if (s->has_tem) s->tem[you] = wall_tem(s, me, p->tem_bc_top, p->tem_top);
      break;
    case 3:
      s->pre[you] = s->vm_pre[me];
      if (s->has_tem) s->tem[you] = wall_tem(s, me, p->tem_bc_left, p->tem_left);
      break;
    case 4:
      s->pre[you] = s->vm_pre[me];
      if (s->has_tem) s->tem[you] = wall_tem(s, me, p->tem_bc_right, p->tem_right);
      break;
    case 5:
    case 6: /* :202-230: bottom corners are always no-slip */
      s->u[you] = -vu;
      s->v[you] = -vv;
      s->pre[you] = wall_pre(s, me, you, p->tem_bc_bottom, p->tem_bottom);
      if (s->has_tem) s->tem[you] = wall_tem(s, me, p->tem_bc_bottom, p->tem_bottom);
      break;
    case 7:
    case 8: /* :235-271 */
      if (p->u_bc_top == 3) {
        s->u[you] = 2.0 * p->u_top - vu;
        s->v[you] = -vv;
      } else {
        s->u[you] = -vu;
        s->v[you] = -vv;
      }
      s->pre[you] = wall_pre(s, me, you, p->tem_bc_top, p->tem_top);
      if (s->has_tem) s->tem[you] = wall_tem(s, me, p->tem_bc_top, p->tem_top);
      break;
    default:
      break;
  }
}

static void extrapolate_diff_test(orc_sim *s, int me) {
  const lssph_params_t *p = &s->p;
  int you = s->n_int + me;
  double vu = s->vm_u[me];
  switch (s->vm_ptype[me]) {
    case 1:
      s->u[you] = (p->u_bc_bottom == 1) ? vu : 2.0 * (sin(ORC_PI * (s->vm_x[me] - p->wl_thick))) - vu;
      break;
    case 2:
      s->u[you] = (p->u_bc_top == 1) ? vu : 2.0 * (-sin(ORC_PI * (s->vm_x[me] - p->wl_thick))) - vu;
      break;
    case 3:
      s->u[you] = (p->u_bc_left == 1)
                      ? vu + ORC_PI * cos(ORC_PI * (s->vm_y[me] - p->wl_thick)) * (s->x[you] - s->vm_x[me])
                      : -vu;
      break;
    case 4:
      s->u[you] = (p->u_bc_right == 1)
                      ? vu + (-ORC_PI) * cos(ORC_PI * (s->vm_y[me] - p->wl_thick)) * (s->x[you] - s->vm_x[me])
                      : -vu;
      break;
    default:
      s->u[you] = -vu;
      break;
  }
}

int32_t orc_ghost(orc_sim *s) {
  if (s->p.periodic || s->n_ext == 0) return 0;
  if (!s->cells_ready) { set_err(s, "ghost before init_cell"); return LSSPH_E_STATE; }
#pragma omp parallel for schedule(static)
  for (int me = 0; me < s->n_ext; ++me) {
    if (s->p.wall_model == 1)
      interpolate_csph(s, me);
    else
      interpolate_b(s, me);
    if (s->p.target_problem == 1)
      extrapolate_diff_test(s, me);
    else
      extrapolate_vm_to_wl(s, me);
  }
  return 0;
}

int32_t orc_get_vm(const orc_sim *s, int32_t *ptype, double *x, double *y,
                   double *u, double *v, double *pre, double *tem) {
  size_t nd = sizeof(double) * (size_t)s->n_ext;
  if (ptype) memcpy(ptype, s->vm_ptype, sizeof(int32_t) * (size_t)s->n_ext);
  if (x) memcpy(x, s->vm_x, nd);
  if (y) memcpy(y, s->vm_y, nd);
  if (u) memcpy(u, s->vm_u, nd);
  if (v) memcpy(v, s->vm_v, nd);
  if (pre) memcpy(pre, s->vm_pre, nd);
  if (tem && s->has_tem) memcpy(tem, s->vm_tem, nd);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Time integration: RK2_mod.f90:42-126, RK4_mod.f90:42-177.                */
/* ------------------------------------------------------------------------ */
static inline double wrap(double x, double L, int periodic) {
  if (!periodic) return x;
  if (x >= L) return x - L;
  if (x < 0.0) return x + L;
  return x;
}

static int after_stage(orc_sim *s) {
  int rc = 0;
  if (s->p.target_problem != 1) rc = orc_update_cell(s); /* :115-117 / :295-297 */
  if (rc) return rc;
  return orc_ghost(s); /* :122-124 / :302-304 */
}

int32_t orc_rk2(orc_sim *s) {
  const lssph_params_t *p = &s->p;
  int n = s->n_int, diff = (p->target_problem == 1), per = p->periodic;
  double c_sq = p->sound_speed * p->sound_speed, rho_ref = p->rho_ref;
  for (int me = 0; me < n; ++me) {
    s->x_old[me] = s->x[me]; s->y_old[me] = s->y[me];
    s->u_old[me] = s->u[me]; s->v_old[me] = s->v[me];
    s->rho_old[me] = s->rho[me];
    if (s->has_tem) s->tem_old[me] = s->tem[me];
  }
  for (int st = 1; st <= 2; ++st) {
    int rc = orc_calculate_rhs(s);
    if (rc) return rc;
    double dt_eff = (st == 1) ? p->dt / 2.0 : p->dt;
#pragma omp parallel for schedule(static)
    for (int me = 0; me < n; ++me) {
      if (!diff) {
        s->x[me] = wrap(s->x_old[me] + s->u[me] * dt_eff, p->len_x, per);
        s->y[me] = wrap(s->y_old[me] + s->v[me] * dt_eff, p->len_y, per);
        s->v[me] = s->v_old[me] + s->rhs_v[me] * dt_eff;
        s->rho[me] = s->rho_old[me] + s->rhs_rho[me] * dt_eff;
      }
      s->u[me] = s->u_old[me] + s->rhs_u[me] * dt_eff;
      if (s->has_tem) s->tem[me] = s->tem_old[me] + s->rhs_tem[me] * dt_eff;
      if (!diff) s->pre[me] = c_sq * (s->rho[me] - rho_ref);
    }
    rc = after_stage(s);
    if (rc) return rc;
  }
  return 0;
}

int32_t orc_rk4(orc_sim *s) {
  const lssph_params_t *p = &s->p;
  int n = s->n_int, diff = (p->target_problem == 1), per = p->periodic, T = s->has_tem;
  double c_sq = p->sound_speed * p->sound_speed, rho_ref = p->rho_ref;
  for (int me = 0; me < n; ++me) {
    s->x_old[me] = s->x[me]; s->y_old[me] = s->y[me];
    s->u_old[me] = s->u[me]; s->v_old[me] = s->v[me];
    s->rho_old[me] = s->rho[me];
    if (T) s->tem_old[me] = s->tem[me];
    s->x_new[me] = s->x[me]; s->y_new[me] = s->y[me];
    s->u_new[me] = s->u[me]; s->v_new[me] = s->v[me];
    s->rho_new[me] = s->rho[me];
    if (T) s->tem_new[me] = s->tem[me];
  }
  for (int st = 1; st <= 4; ++st) {
    int rc = orc_calculate_rhs(s);
    if (rc) return rc;
    double dt1, dt2;
    switch (st) {
      case 1: dt1 = p->dt / 6.0; dt2 = p->dt / 2.0; break;
      case 2: dt1 = p->dt / 3.0; dt2 = p->dt / 2.0; break;
      case 3: dt1 = p->dt / 3.0; dt2 = p->dt; break;
      default: dt1 = p->dt / 6.0; dt2 = 0.0; break;
    }
#pragma omp parallel for schedule(static)
    for (int me = 0; me < n; ++me) {
      s->u_new[me] = s->u_new[me] + s->rhs_u[me] * dt1;
      if (!diff) {
        s->v_new[me] = s->v_new[me] + s->rhs_v[me] * dt1;
        s->x_new[me] = s->x_new[me] + s->u[me] * dt1;
        s->y_new[me] = s->y_new[me] + s->v[me] * dt1;
        s->rho_new[me] = s->rho_new[me] + s->rhs_rho[me] * dt1;
      }
      if (T) s->tem_new[me] = s->tem_new[me] + s->rhs_tem[me] * dt1;
      if (st < 4) {
        if (!diff) {
          s->x[me] = wrap(s->x_old[me] + s->u[me] * dt2, p->len_x, per);
          s->y[me] = wrap(s->y_old[me] + s->v[me] * dt2, p->len_y, per);
          s->v[me] = s->v_old[me] + s->rhs_v[me] * dt2;
          s->rho[me] = s->rho_old[me] + s->rhs_rho[me] * dt2;
        }
        s->u[me] = s->u_old[me] + s->rhs_u[me] * dt2;
        if (T) s->tem[me] = s->tem_old[me] + s->rhs_tem[me] * dt2;
      } else {
        s->u[me] = s->u_new[me];
        if (!diff) {
          s->v[me] = s->v_new[me];
          s->x[me] = wrap(s->x_new[me], p->len_x, per);
          s->y[me] = wrap(s->y_new[me], p->len_y, per);
          s->rho[me] = s->rho_new[me];
        }
        if (T) s->tem[me] = s->tem_new[me];
      }
      if (!diff) s->pre[me] = c_sq * (s->rho[me] - rho_ref);
    }
    rc = after_stage(s);
    if (rc) return rc;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Particle shifting: PST_mod.f90:37-114, shift_interpolate_mod.f90:23-300. */
/* ------------------------------------------------------------------------ */
static void shift_interpolate(orc_sim *s, int me, double coe) {
  const lssph_params_t *p = &s->p;
  int k, q = q_of_model(p->sph_model, &k);
  const int nb = s->has_tem ? 4 : 3;
  double ih[5], a[QMAX], M[QMAX * QMAX], b[QMAX * 4];
  int32_t ipiv[QMAX];
  memset(M, 0, sizeof M);
  memset(b, 0, sizeof b);
  double h = p->h, h_eff_sq = p->h_eff * p->h_eff, m = p->sp_mass, w_ave = p->w_ave;
  double my_x = s->x[me], my_y = s->y[me], my_u = s->u[me], my_v = s->v[me];
  double my_rho = s->rho[me], my_tem = s->has_tem ? s->tem[me] : 0.0;
  double dx = 0.0, dy = 0.0;
  ih[0] = 1.0 / h;
  for (int i = 1; i < k; ++i) ih[i] = ih[i - 1] / h;
  int num = 0, c0 = s->cell_of[me];
  for (int nc = 0; nc < 9; ++nc) {
    int c = s->nbr[9 * c0 + nc];
    if (c < 0) continue;
    double sx = s->nbr_sx[9 * c0 + nc], sy = s->nbr_sy[9 * c0 + nc];
    for (int t = s->cell_start[c]; t < s->cell_start[c + 1]; ++t) {
      int j = s->cell_list[t];
      if (j == me) continue;
      double xj = (s->x[j] + sx) - my_x, yj = (s->y[j] + sy) - my_y;
      double r2 = xj * xj + yj * yj;
      if (r2 > h_eff_sq) continue;
      double r = sqrt(r2);
      double w = orc_kernel_w(p->kernel_type, r, h);
      double dwdr = orc_kernel_dw(p->kernel_type, r, h);
      ++num;
      double vj = m / s->rho[j];
      double vw = vj * w;
      double dwx = dwdr * (-xj) / r * vj;
      double dwy = dwdr * (-yj) / r * vj;
      double wr = w / w_ave;
      double f = 1.0 + 0.2 * (wr * wr * wr * wr);
      dx = dx + f * dwx;
      dy = dy + f * dwy;
      basis_a(q, xj, yj, ih, a);
      double df[4] = {s->u[j] - my_u, s->v[j] - my_v, s->rho[j] - my_rho,
                      s->has_tem ? s->tem[j] - my_tem : 0.0};
      for (int i = 0; i < q; ++i) {
        double vwa = vw * a[i];
        for (int ff = 0; ff < nb; ++ff) b[i + ff * q] += vwa * df[ff];
        for (int rr = 0; rr <= i; ++rr) M[rr + i * q] += vwa * a[rr];
      }
    }
  }
  dx = -coe * dx;
  dy = -coe * dy;
  double rs = sqrt(dx * dx + dy * dy);
  if (rs > 0.2 * h) {
    dx = (0.2 * h) * dx / rs;
    dy = (0.2 * h) * dy / rs;
  }
  s->x_old[me] = s->x[me] + dx;
  s->y_old[me] = s->y[me] + dy;
  for (int i = 0; i < q - 1; ++i)
    for (int r = i + 1; r < q; ++r) M[r + i * q] = M[i + r * q];
  if (num >= q)
    orc_dgesv(q, nb, M, ipiv, b);
  else
    memset(b, 0, sizeof b);
  basis_a(q, dx, dy, ih, a);
  double acc[4] = {0, 0, 0, 0};
  for (int ff = 0; ff < nb; ++ff)
    for (int i = 0; i < q; ++i) acc[ff] += a[i] * b[i + ff * q];
  s->u_old[me] = s->u[me] + acc[0];
  s->v_old[me] = s->v[me] + acc[1];
  s->rho_old[me] = s->rho[me] + acc[2];
  if (s->has_tem) s->tem_old[me] = s->tem[me] + acc[3];
}

int32_t orc_pst(orc_sim *s) {
  const lssph_params_t *p = &s->p;
  if (!s->cells_ready) { set_err(s, "pst before init_cell"); return LSSPH_E_STATE; }
  int n = s->n_int;
  double c_sq = p->sound_speed * p->sound_speed, rho_ref = p->rho_ref;
  double umax = 0.0;
  for (int me = 0; me < n; ++me) {
    double ut = sqrt(s->u[me] * s->u[me] + s->v[me] * s->v[me]);
    umax = fmax(umax, ut);
  }
  double coe = p->pst_c * umax * p->dt * p->h;
#pragma omp parallel for schedule(static)
  for (int me = 0; me < n; ++me) shift_interpolate(s, me, coe);
  for (int me = 0; me < n; ++me) {
    s->x[me] = wrap(s->x_old[me], p->len_x, p->periodic);
    s->y[me] = wrap(s->y_old[me], p->len_y, p->periodic);
    s->u[me] = s->u_old[me];
    s->v[me] = s->v_old[me];
    s->rho[me] = s->rho_old[me];
    if (s->has_tem) s->tem[me] = s->tem_old[me];
    s->pre[me] = c_sq * (s->rho[me] - rho_ref);
  }
  int rc = (p->target_problem != 1) ? orc_update_cell(s) : 0;
  if (rc) return rc;
  return orc_ghost(s);
}

/* check_steady_state_mod.f90:33-67 */
int32_t orc_check_steady_state(orc_sim *s, int32_t *is_steady, double *df) {
  double m = 0.0;
  for (int me = 0; me < s->n_int; ++me) {
    double t = fabs(s->f_prev[me] - s->u[me]);
    m = fmax(t, m);
  }
  if (df) *df = m;
  int st = (m < s->p.threshold) ? 1 : 0;
  if (is_steady) *is_steady = st;
  if (!st)
    for (int me = 0; me < s->n_int; ++me) s->f_prev[me] = s->u[me];
  return 0;
}

/* Neighbor sets in summation order (LSSPH_A_mod.f90:102-124). */
int32_t orc_neighbors(const orc_sim *s, int32_t *counts, int32_t *list,
                      int64_t capacity, int64_t *total) {
  int64_t w = 0;
  double h_eff_sq = s->p.h_eff * s->p.h_eff;
  for (int me = 0; me < s->n_int; ++me) {
    int cnt = 0, c0 = s->cell_of[me];
    for (int nc = 0; nc < 9; ++nc) {
      int c = s->nbr[9 * c0 + nc];
      if (c < 0) continue;
      double sx = s->nbr_sx[9 * c0 + nc], sy = s->nbr_sy[9 * c0 + nc];
      for (int t = s->cell_start[c]; t < s->cell_start[c + 1]; ++t) {
        int j = s->cell_list[t];
        if (j == me) continue;
        double xj = (s->x[j] + sx) - s->x[me], yj = (s->y[j] + sy) - s->y[me];
        if (xj * xj + yj * yj > h_eff_sq) continue;
        if (w < capacity) list[w] = j;
        ++w;
        ++cnt;
      }
    }
    counts[me] = cnt;
  }
  if (total) *total = w;
  return (w > capacity) ? LSSPH_E_CAPACITY : 0;
}
