/*
 * lssph.h -- C ABI of the MI355X (gfx950) LS-SPH hot path.
 *
 * One handle owns the device-resident particle state of one simulation, or of
 * an ensemble of independent simulations that share the model options, plus
 * its cell lists, ghost-wall markers, HIP stream and captured step graphs.
 * Every entry point returns 0 on success and a negative LSSPH_E* code on
 * failure; lssph_last_error() then holds a message.  Calls on one handle are
 * asynchronous on the handle's stream except those that return host data
 * (download, check_steady_state, neighbors, stats).  A handle is not
 * reentrant: one host thread at a time.
 *
 * Array conventions: plain host pointers, element 0 = Fortran element 1.
 * Particle arrays have n_total = sum(n_int) + sum(n_ext) entries ordered as
 * the reference orders them for one box (source/setup/setup_closed_box_mod.f90:
 * 185-193: all fluid particles first, then all ghost-wall particles); for an
 * ensemble the fluid blocks of the instances come first, in instance order,
 * followed by their wall blocks in instance order.  Virtual marker k pairs
 * with wall particle sum(n_int) + k (source/setup/init_virtual_markers_mod.f90:
 * 61-64).
 */
#ifndef LSSPH_H
#define LSSPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSSPH_OK 0
#define LSSPH_E_ARG -1       /* invalid argument or unsupported option        */
#define LSSPH_E_HIP -2       /* HIP runtime error (message has the HIP text)  */
#define LSSPH_E_STATE -3     /* call out of order (e.g. step before upload)   */
#define LSSPH_E_DOMAIN -4    /* a fluid particle left the cell grid, or moved
                                farther than the skin allows between builds;
                                its instance stops advancing at that pass
                                until the next lssph_upload                  */
#define LSSPH_E_NOMEM -5     /* device allocation failed                      */
#define LSSPH_E_CAPACITY -6  /* caller buffer too small                       */

/* precision codes */
#define LSSPH_FP64 0   /* all FP64 (reference arithmetic)                     */
#define LSSPH_MIXED 1  /* FP32 pair sums, FP64 factorisation + refinement     */

/*
 * Numeric parameters and model options of one instance: the C-interoperable
 * part of Param_type (source/core/global_types.f90:22-73) plus the config.h
 * model switches.  Derived fields (dx ... dt) are filled by the host exactly
 * as init_param does (source/setup/init_param_mod.f90:92-199).  A Fortran
 * caller mirrors it with a bind(C) derived type of the same field order.
 */
typedef struct lssph_params {
  /* domain and resolution */
  double len_x, len_y;      /* LEN_X, LEN_Y                                   */
  double dx;                /* len_x / num_x                                  */
  double sp_mass;           /* rho_ref * dx**2                                */
  double h;                 /* coe_h * dx                                     */
  double h_eff;             /* 2h (3h for the quintic spline)                 */
  double w_ave;             /* W(dx, h)                                       */
  double wl_thick;          /* num_wl * dx (0 when periodic)                  */
  /* physics */
  double rho_ref, vis_ref, sound_speed, thermal_dif;
  double alpha_ref, gravity, tem_ave;
  double tem_top, tem_bottom, tem_left, tem_right;
  double u_top;
  double zeta_rsst, xi_vim, coe_delta_sph, pst_c;
  double dt;
  double threshold;         /* steady-state threshold (diffusion test)        */
  /* extensions (reference value in brackets) */
  double diff_coef;         /* factor on the Laplacian in the diffusion RHS [1] */
  double diff_source;       /* factor on the manufactured source term       [1] */
  double cond_max;          /* mixed precision: a particle whose 5x5 moment
                               matrix has a 1-norm condition number
                               ||M||_1 ||M^-1||_1 (exact, from the LDL^T
                               factors) above this is redone in full FP64,
                               sums included (<= 0 selects 1e4)             */
  /* model options (config.h macros) */
  int32_t target_problem;   /* TARGET_PROBLEM 1..4                            */
  int32_t kernel_type;      /* KERNEL_TYPE 1..5                               */
  int32_t sph_model;        /* SPH_MODEL 1..6: 1-2 classical SPH (sum /
                               difference pressure gradient), 3-6 LS-SPH type
                               A with q = 5, 9, 14, 20; LSSPH_MIXED applies to
                               3, the other models compute in FP64           */
  int32_t wall_model;       /* WALL_MODEL 1..3                                */
  int32_t rk;               /* RK 2 or 4                                      */
  int32_t u_bc_top, u_bc_bottom, u_bc_left, u_bc_right;       /* U_BOUNDARY_*   */
  int32_t tem_bc_top, tem_bc_bottom, tem_bc_left, tem_bc_right; /* TEM_BOUNDARY_* */
  int32_t num_x, num_y, num_wl;
  int32_t periodic;         /* 0: closed box with ghost walls (reference)
                               1: doubly periodic box, no walls (extension)   */
  int32_t precision;        /* LSSPH_FP64 or LSSPH_MIXED                      */
} lssph_params_t;

/* Counters the device keeps; read with lssph_stats(). */
typedef struct lssph_stats {
  int64_t fp64_fallbacks;   /* mixed path: particle passes (operator or
                               shift) redone in full FP64 after the
                               condition check                              */
  int64_t list_overflows;   /* particles whose neighbor list exceeded the
                               on-chip list and took the direct path          */
  int32_t max_neighbors;    /* largest neighbor count seen                    */
  int32_t max_cell_count;   /* most particles in one cell                     */
  int32_t escaped;          /* fluid particles found outside the cell grid    */
  int32_t pad;
} lssph_stats_t;

typedef struct lssph_handle *lssph_handle_t;

/* Replaces: allocation in setup_closed_box (setup_closed_box_mod.f90:142-183),
   init_virtual_markers (init_virtual_markers_mod.f90:37-48) and init_cell
   (init_cell_mod.f90:46-60,99-100).  params[n_inst], n_int[n_inst],
   n_ext[n_inst]; options must agree across instances. */
int32_t lssph_create(const lssph_params_t *params, int32_t n_inst,
                     const int32_t *n_int, const int32_t *n_ext,
                     lssph_handle_t *out);
int32_t lssph_destroy(lssph_handle_t h);

/* Message of the last failure on h (or of the last failed lssph_create when
   h is NULL).  Valid until the next call on h. */
const char *lssph_last_error(lssph_handle_t h);

/* Host -> device copy of SP (global_types.f90:78-85).  tem may be NULL unless
   target_problem == 4.  Also derives the virtual markers from the wall
   positions (init_virtual_markers_mod.f90:61-137). */
int32_t lssph_upload(lssph_handle_t h, const int32_t *ptype, const double *x,
                     const double *y, const double *u, const double *v,
                     const double *rho, const double *pre, const double *tem);

/* Device -> host copy of SP in the upload order (for write_data,
   write_data_mod.f90:43-52).  Synchronises. */
int32_t lssph_download(lssph_handle_t h, int32_t *ptype, double *x, double *y,
                       double *u, double *v, double *rho, double *pre,
                       double *tem);

/* Replaces init_cell + update_cell (init_cell_mod.f90:23-184,
   update_cell_mod.f90:21-97): builds the cell grid, the wall and marker cell
   indices, and bins every particle. */
int32_t lssph_init_cell(lssph_handle_t h);

/* Replaces update_cell (update_cell_mod.f90:21-97). */
int32_t lssph_update_cell(lssph_handle_t h);

/* Replaces ghost (ghost_mod.f90:43-78): marker interpolation + extrapolation
   to the paired wall particles.  No-op for a periodic box. */
int32_t lssph_ghost(lssph_handle_t h);

/* Replaces calculate_RHS (calculate_RHS_mod.f90:29-52) / LSSPH_A
   (LSSPH_A_mod.f90:26-271); results read back with lssph_download_rhs. */
int32_t lssph_calculate_rhs(lssph_handle_t h);

/* RHS_u, RHS_v, RHS_rho, RHS_tem of the fluid particles (sum(n_int) each, in
   upload order).  rhs_tem may be NULL.  Synchronises. */
int32_t lssph_download_rhs(lssph_handle_t h, double *rhs_u, double *rhs_v,
                           double *rhs_rho, double *rhs_tem);

/* Replace RK2 (RK2_mod.f90:30-128) and RK4 (RK4_mod.f90:30-179): one time
   step with the per-stage update_cell and ghost. */
int32_t lssph_rk2(lssph_handle_t h);
int32_t lssph_rk4(lssph_handle_t h);

/* Replaces PST (PST_mod.f90:28-116) incl. shift_interpolate
   (shift_interpolate_mod.f90:23-300), its update_cell and ghost. */
int32_t lssph_pst(lssph_handle_t h);

/* nsteps iterations of the time loop body of main.f90:119-134 (RK2 or RK4
   as params.rk says, then PST unless target_problem == 1), replayed from a
   captured HIP graph. */
int32_t lssph_step(lssph_handle_t h, int32_t nsteps);

/* Replaces check_steady_state (check_steady_state_mod.f90:23-69): computes
   max|f_previous - u| per instance; for instances not yet steady copies u to
   f_previous.  is_steady[n_inst], df_max[n_inst] (either may be NULL).
   Synchronises. */
int32_t lssph_check_steady_state(lssph_handle_t h, int32_t *is_steady,
                                 double *df_max);

/* Virtual-marker data as write_param writes VM.dat (write_param_mod.f90:
   147-157); sum(n_ext) entries each, tem may be NULL.  Synchronises. */
int32_t lssph_download_vm(lssph_handle_t h, int32_t *ptype, double *x,
                          double *y, double *u, double *v, double *pre,
                          double *tem);

/* cell%idx_SP (1-based cell index within the particle's own instance grid),
   as write_param writes cell_idx.dat (write_param_mod.f90:133-136). */
int32_t lssph_download_cell_index(lssph_handle_t h, int32_t *idx_sp);

/* Validation: neighbor sets N(i) = { j != i in the 3x3 cell stencil with
   r_ij^2 <= h_eff^2 } of every fluid particle (LSSPH_A_mod.f90:102-124), in
   summation order, as 0-based upload-order indices.  counts[sum(n_int)];
   list receives the concatenation (capacity entries available); *total gets
   the number written.  With list NULL or capacity below the total, counts
   and *total are still filled and the call returns LSSPH_E_CAPACITY (the
   sizing call).  Synchronises. */
int32_t lssph_neighbors(lssph_handle_t h, int32_t *counts, int32_t *list,
                        int64_t capacity, int64_t *total);

/* Validation of the pair kernels' own neighbor lists (no reference
   counterpart; the sets are those of LSSPH_A_mod.f90:102-124).  Runs the
   candidate scan and the list walk of the LS operator and shifting kernels
   (same code, launch shape and on-chip layout, under the current
   lssph_set_scan / lssph_set_tuning settings) and records, per fluid particle
   in upload order, the neighbors the walk visits, in its order:
   list[i * cap + k], k < counts[i], holds 0-based upload indices (-1 - j
   would mark a visit beyond h_eff); path[i] is the scan path taken: 0
   per-particle scan, 1 wave-staged scan row by row, 2 wave-staged scan with
   positions kept on chip, 3 list overflow (direct stencil walk).  Returns
   LSSPH_E_CAPACITY when some counts[i] > cap (counts are still filled).
   Not for SPH_MODEL 6, whose kernels gather per wavefront.  Synchronises. */
int32_t lssph_pair_lists(lssph_handle_t h, int32_t cap, int32_t *counts,
                         int32_t *list, int32_t *path);

/* Type-B least-squares fit of one field at n probe points of instance inst,
   as the reference's analysis interpolates wall values and gradients
   (analysis/common/lssph_b.py:21-144, used for the Nusselt number by
   analysis/benchmarks/boussinesq_conv.py:80-108): all particles, fluid and
   wall, within h_eff of a probe; q = 3, 6 or 10 basis terms with a constant.
   x, y are in the upload coordinates (walls from 0 to wl_thick); field 0 u,
   1 v, 2 rho, 3 pre, 4 tem.  out[3k], out[3k+1], out[3k+2] = value, d/dx,
   d/dy at probe k (NaN with fewer than q particles or a singular fit).
   FP64 in either precision mode.  Synchronises. */
int32_t lssph_probe(lssph_handle_t h, int32_t inst, int32_t field, int32_t q,
                    int32_t n, const double *x, const double *y, double *out);

/* lssph_probe without any allocation: the same fit and arguments, with the
   probe coordinates and results staged in a fixed device buffer of the
   library (one per GPU, held until the process ends) instead of buffers
   allocated per call.  No counterpart in the reference; it is for callers
   that cannot allocate device memory (tight monitoring loops inside a time
   loop, harnesses that forbid hipMalloc).  At most 16384 probes per call
   (LSSPH_E_ARG beyond).  Calls on one GPU share the buffer: calls that
   overlap in time on one GPU (several host threads or handles) race.
   Synchronises. */
int32_t lssph_probe_static(lssph_handle_t h, int32_t inst, int32_t field,
                           int32_t q, int32_t n, const double *x,
                           const double *y, double *out);

/* Ensembles: isolation of diverged instances (no reference counterpart: the
   reference runs one box and, when its scheme diverges, indexes past its cell
   arrays).  on = 1: a fluid particle that leaves its instance's cell grid,
   or moves farther than the cell-list skin allows between builds (a
   non-finite position does both), retires its instance instead of failing
   the handle.  From the next pass on, the pair kernels and ghost walls leave
   that instance's records as they are; the other instances run on, and
   their results are the same as without the retired one.  0 (the default):
   such a move is an LSSPH_E_DOMAIN error of the handle, reported by the next
   call that returns host data, and the instance is frozen in the same way
   (a diverged run stops costing time) until the next lssph_upload.  Turning it on
   clears every instance's status and restarts the step count; turning it
   off leaves instances already retired frozen (lssph_upload revives them
   with new data).  A too small skin (lssph_set_skin) for an instance's
   velocities also retires it: its moves outrun the cell list. */
int32_t lssph_set_isolation(lssph_handle_t h, int32_t on);

/* retired[n_inst]: 0 while instance k runs, else the time step (counted
   from lssph_set_isolation, one per lssph_rk2/lssph_rk4 call or lssph_step
   iteration) during which it was retired; reason[n_inst] (may be NULL): the
   first cause seen, LSSPH_RETIRED_SKIN (a move beyond the cell-list skin:
   the instance's neighbor sets could no longer be guaranteed -- a diverging
   run, or a skin too small for the instance's velocities) or
   LSSPH_RETIRED_ESCAPED (a position outside the grid or not finite).
   Synchronises. */
#define LSSPH_RETIRED_SKIN 1
#define LSSPH_RETIRED_ESCAPED 2
int32_t lssph_instance_status(lssph_handle_t h, int32_t *retired,
                              int32_t *reason);

/* Device counters (see lssph_stats_t).  Synchronises. */
int32_t lssph_stats(lssph_handle_t h, lssph_stats_t *out);
int32_t lssph_reset_stats(lssph_handle_t h);

/* Waits for all work queued on the handle's stream. */
int32_t lssph_synchronize(lssph_handle_t h);

/* Tuning knobs (set by hand, or picked per workload by the autotuner in
   ls-sph-benchmarks_amd/lssph/tune.py): threads per block of the pair kernels,
   on-chip neighbor-list length (entries per particle, the particle itself
   included; a longer list walks the candidates instead; lssph_create sizes
   it from pi (h_eff/dx)^2), and how many time steps (lssph_step, or
   lssph_rk2/lssph_rk4 calls) pass between physical re-sorts of the particle
   records (default 50; negative: never).  0 leaves a value unchanged.  The
   resulting pair must keep (list_len + 1) * block * 4 bytes of LDS within
   64 KiB; a rejected call changes nothing. */
int32_t lssph_set_tuning(lssph_handle_t h, int32_t block, int32_t list_len,
                         int32_t reorder_interval);

/* Cell-list skin as a fraction of h_eff (default 0.08).  The library has no
   counterpart in the reference, whose update_cell (update_cell_mod.f90:
   21-97) rebuilds h_eff-wide cells after every RK stage and shift: here the
   cells are h_eff * (1 + skin) wide and one list serves the RK stages of a
   time step (it is rebuilt after the shift, or after rk + 1 stages); every
   pass still tests r^2 <= h_eff^2 at the current positions, so the
   neighbor sets are unchanged.  A particle moving more than skin * h_eff / 2
   between builds makes the next synchronising call fail with
   LSSPH_E_DOMAIN.  0 rebuilds the cells after every pass, as the reference
   does.  Call before lssph_upload. */
int32_t lssph_set_skin(lssph_handle_t h, double skin);

/* Cell columns per stencil reach in x, 1 to 4 (default 2; no reference
   counterpart: the reference's init_cell_mod.f90:36-41 grid is square, as
   with 1).  The internal cells stay h_eff * (1 + skin) tall and become
   h_eff * (1 + skin) / columns wide; a particle's stencil is the
   2 columns + 1 cells of the rows above, at and below its cell, which holds
   about (2 columns + 1) / (3 columns) of the candidates of a square 3x3
   stencil.  The neighbor sets do not change; the order of the candidates
   within a stencil row does, so the sums round differently (as with another
   skin).  Call before lssph_upload. */
int32_t lssph_set_cell_split(lssph_handle_t h, int32_t columns);

/* Candidate scan of the pair kernels (no reference counterpart).  1 (the
   default): a full wavefront whose particles lie in one cell row loads its
   candidates' positions once into on-chip memory and scans them there;
   other wavefronts scan per particle.  0: every wavefront scans per
   particle.  Both give the same neighbor lists in the same order, so the
   results are bitwise identical. */
int32_t lssph_set_scan(lssph_handle_t h, int32_t staged);

/* Particles per wavefront of the SPH_MODEL 6 kernels (q = 20), a tuning
   knob with no reference counterpart.  1 or 2: a group of 64 / ppw lanes
   gathers, sums and factors one particle's 20 x 20 system in on-chip
   memory; 3: a group of 20 lanes, one matrix row per lane, factors it with
   the rows in registers.  0 restores the default (3).  All values give the
   same results bit for bit; the other models (one particle per lane) ignore
   it. */
int32_t lssph_set_wave_particles(lssph_handle_t h, int32_t ppw);

/* The HIP stream (hipStream_t) the handle launches on, for event timing. */
void *lssph_stream(lssph_handle_t h);

/* Per-kernel timing.  While enabled, eager calls (lssph_rk2/rk4/pst/ghost/
   update_cell, not the graphs of lssph_step) record a HIP event pair on the
   handle's stream around every launch of each kernel group;
   lssph_kernel_times() synchronizes, returns the accumulated milliseconds and
   launch counts per group, and clears them.  No reference counterpart (the
   reference times only its whole loop, main.f90:110,162). */
typedef struct lssph_kernel_times {
  double lsa_ms;     /* fused LS type-A operator + RK stage update          */
  double pst_ms;     /* particle shifting                                   */
  double cells_ms;   /* cell-list build (5 kernels per build)               */
  double ghost_ms;   /* ghost-wall markers                                  */
  double other_ms;   /* max-speed reduction                                 */
  int64_t lsa_launches, pst_launches, cells_launches, ghost_launches, other_launches;
} lssph_kernel_times_t;
int32_t lssph_profile(lssph_handle_t h, int32_t enable);
int32_t lssph_kernel_times(lssph_handle_t h, lssph_kernel_times_t *out);

#ifdef __cplusplus
}
#endif

#endif /* LSSPH_H */
