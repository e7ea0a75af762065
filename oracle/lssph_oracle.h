/*
 * lssph_oracle.h -- CPU oracle of the LS-SPH hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Fortran loops (shobuzako-kensuke/LS-SPH-Benchmarks, source/...) used as
 * the checker for the HIP library.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never calls it.
 *
 * One orc_sim holds one instance (one box).  Arrays use the reference's
 * particle order (fluid first, then ghost walls) with 0-based indices.
 */
#ifndef LSSPH_ORACLE_H
#define LSSPH_ORACLE_H

#include <stdint.h>
#include "../include/lssph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

orc_sim *orc_create(const lssph_params_t *p, int32_t n_int, int32_t n_ext);
void orc_destroy(orc_sim *s);
const char *orc_error(const orc_sim *s);
void orc_set_threads(int32_t n);

int32_t orc_set_state(orc_sim *s, const int32_t *ptype, const double *x,
                      const double *y, const double *u, const double *v,
                      const double *rho, const double *pre, const double *tem);
int32_t orc_get_state(const orc_sim *s, int32_t *ptype, double *x, double *y,
                      double *u, double *v, double *rho, double *pre,
                      double *tem);
int32_t orc_init_vm(orc_sim *s);
int32_t orc_init_cell(orc_sim *s);
int32_t orc_update_cell(orc_sim *s);
int32_t orc_ghost(orc_sim *s);
int32_t orc_calculate_rhs(orc_sim *s);
int32_t orc_get_rhs(const orc_sim *s, double *ru, double *rv, double *rrho,
                    double *rtem);
int32_t orc_rk2(orc_sim *s);
int32_t orc_rk4(orc_sim *s);
int32_t orc_pst(orc_sim *s);
int32_t orc_check_steady_state(orc_sim *s, int32_t *is_steady, double *df);
int32_t orc_get_vm(const orc_sim *s, int32_t *ptype, double *x, double *y,
                   double *u, double *v, double *pre, double *tem);
int32_t orc_get_cell_index(const orc_sim *s, int32_t *idx_sp);
int32_t orc_neighbors(const orc_sim *s, int32_t *counts, int32_t *list,
                      int64_t capacity, int64_t *total);

/* Kernel function and its radial derivative (kernel_functions_mod.f90). */
double orc_kernel_w(int32_t kernel_type, double r, double h);
double orc_kernel_dw(int32_t kernel_type, double r, double h);

/* LAPACK DGESV semantics on column-major A (n x n) and B (n x nrhs). */
int32_t orc_dgesv(int32_t n, int32_t nrhs, double *a, int32_t *ipiv,
                  double *b);

#ifdef __cplusplus
}
#endif

#endif
