// ls_common.hpp -- pieces shared by the LS-SPH pair operators: field layout,
// LS sums and bases, the mixed-precision solve, kernel arguments, and the
// RHS + Runge-Kutta stage epilogue.
#pragma once

#include <type_traits>

#include "pair_common.hpp"

namespace lssph {

// ---------------------------------------------------------------------------
// Field layout of the LS type-A right-hand sides.
//   CLS 0 (diffusion test): u
//   CLS 1 (flow):           u, v, pre [, rho when delta-SPH is on]
//   CLS 2 (Boussinesq):     u, v, pre, tem [, rho]
// rho's column only feeds the delta-SPH term (LSSPH_A_mod.f90:253), so it is
// carried only when coe_delta_sph != 0.
// ---------------------------------------------------------------------------
template <int CLS, bool DELTA>
struct Layout {
  static constexpr int U = 0, V = 1, P = 2, T = 3;
  static constexpr int NF = CLS == 0 ? 1 : (CLS == 1 ? 3 : 4) + (DELTA ? 1 : 0);
  static constexpr int RHO = NF - 1;  // valid when DELTA
};

// The mixed path's differences: u, v and pre from the FP32 neighbor fields
// (g of the neighbor, own of the particle), tem and the delta-SPH rho from
// the FP64 records (the neighbor fields hold four values).
template <int CLS, bool DELTA>
__device__ __forceinline__ void field_diffs32(const float4 &g, const float4 &own, const Rec &nb, const Rec &me,
                                              float (&df)[Layout<CLS, DELTA>::NF]) {
  using L = Layout<CLS, DELTA>;
  df[L::U] = g.x - own.x;
  if constexpr (CLS >= 1) {
    df[L::V] = g.y - own.y;
    df[L::P] = g.z - own.z;
    if constexpr (CLS == 2) df[L::T] = float(nb.tem - me.tem);
    if constexpr (DELTA) df[L::RHO] = float(nb.rho - me.rho);
  }
}

template <int CLS, bool DELTA, typename R>
__device__ __forceinline__ void field_diffs(const Rec &nb, const Rec &me, R (&df)[Layout<CLS, DELTA>::NF]) {
  using L = Layout<CLS, DELTA>;
  df[L::U] = R(nb.u - me.u);
  if constexpr (CLS >= 1) {
    df[L::V] = R(nb.v - me.v);
    df[L::P] = R(nb.pre - me.pre);
    if constexpr (CLS == 2) df[L::T] = R(nb.tem - me.tem);
    if constexpr (DELTA) df[L::RHO] = R(nb.rho - me.rho);
  }
}

// Moment matrix (upper triangle, index i*(i+1)/2 + r for r <= i) and the
// right-hand sides of a q-term LS fit: q = 5, or q = 2 for the shifting of
// the classical SPH models (shift_interpolate_mod.f90:43-44).
template <typename R, int NF, int Q = 5>
struct LsSums {
  static constexpr int NM = Q * (Q + 1) / 2;
  R M[NM];
  R b[NF][Q];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < NM; ++i) M[i] = R(0);
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int i = 0; i < Q; ++i) b[f][i] = R(0);
  }
  // LSSPH_A_mod.f90:174-190
  __device__ __forceinline__ void add(R vw, const R (&a)[Q], const R (&df)[NF]) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const R vwa = vw * a[i];
#pragma unroll
      for (int f = 0; f < NF; ++f) b[f][i] += vwa * df[f];
#pragma unroll
      for (int r = 0; r <= i; ++r) M[i * (i + 1) / 2 + r] += vwa * a[r];
    }
  }
  // Mixed path, q = 5, with the basis of basis5s, [x, y, x^2, y^2, xy]: the
  // moment matrix has 12 distinct sums, M(2,1) = M(4,0), M(3,0) = M(4,1) and
  // M(3,2) = M(4,4), so add_reduced skips those three and complete() copies
  // them after the loop (the results differ from direct accumulation by
  // FP32 rounding of the sums only).
  __device__ __forceinline__ void add_reduced(R vw, const R (&a)[Q], const R (&df)[NF]) {
    static_assert(Q == 5, "q = 5 basis");
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const R vwa = vw * a[i];
#pragma unroll
      for (int f = 0; f < NF; ++f) b[f][i] += vwa * df[f];
#pragma unroll
      for (int r = 0; r <= i; ++r) {
        if ((i == 2 && r == 1) || (i == 3 && r == 0) || (i == 3 && r == 2)) continue;
        M[i * (i + 1) / 2 + r] += vwa * a[r];
      }
    }
  }
  __device__ __forceinline__ void complete() {
    M[4] = M[10];  // (2,1) from (4,0)
    M[6] = M[11];  // (3,0) from (4,1)
    M[8] = M[14];  // (3,2) from (4,4)
  }
};

// Taylor basis a = [x/h, y/h, x^2/2h^2, y^2/2h^2, xy/h^2] (LSSPH_A_mod.f90:
// 140-145); ih2 = (1/h)/h as the reference's inv_h(2).
__device__ __forceinline__ void basis5(double x, double y, double ih, double ih2, double (&a)[5]) {
  a[0] = x * ih;
  a[1] = y * ih;
  a[2] = 0.5 * (x * x) * ih2;
  a[3] = 0.5 * (y * y) * ih2;
  a[4] = x * y * ih2;
}

// The q-term basis: q = 2 keeps the first-order terms (the shifting of the
// classical models), q = 5 is basis5, q = 9 and 14 add the 3rd- and
// 4th-order terms (LSSPH_A_mod.f90:148-162, shift_interpolate_mod.f90:
// 159-172).
template <int Q>
__device__ __forceinline__ void basis_q(double x, double y, const InstDev &I, double (&a)[Q]) {
  static_assert(Q == 2 || Q == 5 || Q == 9 || Q == 14, "q = 2, 5, 9 or 14");
  if constexpr (Q == 2) {
    a[0] = x * I.inv_h;
    a[1] = y * I.inv_h;
  } else {
    a[0] = x * I.inv_h;
    a[1] = y * I.inv_h;
    a[2] = 0.5 * (x * x) * I.inv_h2;
    a[3] = 0.5 * (y * y) * I.inv_h2;
    a[4] = x * y * I.inv_h2;
  }
  if constexpr (Q >= 9) {
    a[5] = 1.0 / 6.0 * (x * x * x) * I.inv_h3;
    a[6] = 1.0 / 6.0 * (y * y * y) * I.inv_h3;
    a[7] = 0.5 * (x * x) * y * I.inv_h3;
    a[8] = 0.5 * (y * y) * x * I.inv_h3;
  }
  if constexpr (Q >= 14) {
    a[9] = 1.0 / 24.0 * (x * x * x * x) * I.inv_h4;
    a[10] = 1.0 / 24.0 * (y * y * y * y) * I.inv_h4;
    a[11] = 1.0 / 6.0 * (x * x * x) * y * I.inv_h4;
    a[12] = 1.0 / 6.0 * (y * y * y) * x * I.inv_h4;
    a[13] = 1.0 / 4.0 * (x * x) * (y * y) * I.inv_h4;
  }
}

// The basis from offsets already scaled by 1/h (mixed precision), without
// the factors 1/2 of the squares: [x, y, x^2, y^2, xy].  Every sum of M and
// b is then the one of basis5 times an exact power of two (S M S and S b,
// S = diag(1, 1, 2, 2, 1)), and LDL^T without pivoting commutes with such a
// scaling, so ldlt_solve_refine returns bitwise the coefficients of the
// halved basis after multiplying entries 2 and 3 by 2; two multiplications
// per neighbor are saved.
__device__ __forceinline__ void basis5s(float xs, float ys, float (&a)[5]) {
  a[0] = xs;
  a[1] = ys;
  a[2] = xs * xs;
  a[3] = ys * ys;
  a[4] = xs * ys;
}

// Coefficients of the basis5s sums -> coefficients of the basis5 terms.
template <int NF>
__device__ __forceinline__ void unscale_squares(double (&X)[NF][5]) {
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    X[f][2] *= 2.0;
    X[f][3] *= 2.0;
  }
}

template <typename R, int NF, int Q>
__device__ __forceinline__ void unpack(const LsSums<R, NF, Q> &S, double (&A)[Q][Q], double (&B)[NF][Q]) {
#pragma unroll
  for (int i = 0; i < Q; ++i)
#pragma unroll
    for (int r = 0; r <= i; ++r) {
      A[r][i] = double(S.M[i * (i + 1) / 2 + r]);
      A[i][r] = A[r][i];
    }
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int i = 0; i < Q; ++i) B[f][i] = double(S.b[f][i]);
}

// Symmetric 5x5 entry (i, c) of packed upper-triangle sums.
template <typename R>
__device__ __forceinline__ double sym5(const R (&M)[15], int i, int c) {
  return i >= c ? double(M[i * (i + 1) / 2 + c]) : double(M[c * (c + 1) / 2 + i]);
}

// FP64 LDL^T of the symmetric 5x5 moment matrix held as FP32 sums (of the
// basis5s basis), with one step of iterative refinement per right-hand side
// (x += M^-1 (b - M x), the residual formed in FP64 against the FP32 sums).
// The right-hand sides are processed one at a time and the sums stay FP32 in
// registers, which keeps the solve's register footprint near that of the
// pair loop.  X holds the coefficients of the basis5 terms.
//
// Returns false -- the caller then redoes the particle in full FP64 -- when
// a pivot is not positive or when the 1-norm condition number of the basis5
// system, kappa_1 = ||M||_1 ||M^-1||_1, exceeds cond_max (or is not finite).
// kappa_1 is exact here, not an estimate that can undershoot: M^-1 =
// L^-T D^-1 L^-1 is formed from the factors (about 130 FP64 operations per
// particle; ||M||_1 is taken on the FP32 sums) -- unless a diagonal-dominance
// bound already puts kappa_1 below half of cond_max (about 40 operations;
// the usual case for interior particles).  The basis5 system is
// D M D with D = diag(1, 1, 1/2, 1/2, 1),
// so its entries and those of its inverse are the basis5s ones times
// D_i D_j and 1 / (D_i D_j).  (The pivot ratio max(d) / min(d) that this
// replaces is a lower bound of cond_2 and misses correlated basis columns:
// on one-sided supports it falls short of kappa_1 by up to 12x.)
__device__ __forceinline__ double basis5_weight(int i) { return (i == 2 || i == 3) ? 0.5 : 1.0; }

// 1 / d for a positive normal d: the hardware reciprocal and two Newton
// steps (within an ulp; the correctly rounded division costs about twice as
// many instructions).  Used by the mixed path's factorization only.
__device__ __forceinline__ double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

template <int NF>
__device__ __forceinline__ bool ldlt_solve_refine(const LsSums<float, NF> &S, double (&X)[NF][5],
                                                  double cond_max) {
  double L[5][5], d[5], dinv[5];
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    double dj = sym5(S.M, j, j);
#pragma unroll
    for (int k = 0; k < j; ++k) dj -= L[j][k] * L[j][k] * d[k];
    ok = ok && (dj > 0.0);
    d[j] = dj;
    dinv[j] = recip(dj);
#pragma unroll
    for (int i = j + 1; i < 5; ++i) {
      double t = sym5(S.M, i, j);
#pragma unroll
      for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k] * d[k];
      L[i][j] = t * dinv[j];
    }
  }
  if (!ok) return false;
  // ||M||_1 on the FP32 sums
  float nrm_m = 0.0f;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    float col = 0.0f;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      const float e = i >= j ? S.M[i * (i + 1) / 2 + j] : S.M[j * (j + 1) / 2 + i];
      col += fabsf(e) * float(basis5_weight(i));
    }
    nrm_m = fmaxf(nrm_m, col * float(basis5_weight(j)));
  }
  // A sufficient test first.  When the symmetric basis5 system B is
  // strictly diagonally dominant, g = min_i (B_ii - sum_j!=i |B_ij|) > 0,
  // Varah's bound (Linear Algebra Appl. 11, 1975) gives ||B^-1||_inf <= 1 / g,
  // and ||B^-1||_1 = ||B^-1||_inf for symmetric B, so kappa_1 <= ||B||_1 / g.
  // A system whose bound is at most half of cond_max has kappa_1 < cond_max
  // (the factor 2 covers the rounding of both sides), so it passes the
  // exact test below, which then need not run: the decisions, and so the
  // results, are those of the exact test alone.
  double g = 1.0e300, bmax = 0.0;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    double off = 0.0;
#pragma unroll
    for (int c = 0; c < 5; ++c)
      if (c != i) off += fabs(sym5(S.M, i, c)) * (basis5_weight(i) * basis5_weight(c));
    const double bii = sym5(S.M, i, i) * (basis5_weight(i) * basis5_weight(i));
    g = fmin(g, bii - off);
    bmax = fmax(bmax, bii);
  }
  const bool bounded = g >= 1.0e-3 * bmax && double(nrm_m) <= 0.5 * cond_max * g;
  if (!bounded) {
    // ||M^-1||_1 one column at a time: column j solves M y = e_j with the
    // factors (L y' = e_j, y' /= d, L^T y = y'), so only five doubles of the
    // inverse are live at once
    double nrm_inv = 0.0;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      double y[5];
#pragma unroll
      for (int i = 0; i < 5; ++i) {
        double t = (i == j) ? 1.0 : 0.0;
#pragma unroll
        for (int k = j; k < i; ++k) t -= L[i][k] * y[k];
        y[i] = (i < j) ? 0.0 : t;
      }
#pragma unroll
      for (int i = j; i < 5; ++i) y[i] *= dinv[i];
#pragma unroll
      for (int i = 3; i >= 0; --i)
#pragma unroll
        for (int k = i + 1; k < 5; ++k) y[i] -= L[k][i] * y[k];
      double col = 0.0;
#pragma unroll
      for (int i = 0; i < 5; ++i) col += fabs(y[i]) / basis5_weight(i);
      nrm_inv = fmax(nrm_inv, col / basis5_weight(j));
    }
    if (!(double(nrm_m) * nrm_inv <= cond_max)) return false;
  }
  auto solve = [&](double (&y)[5]) {
#pragma unroll
    for (int i = 1; i < 5; ++i)
#pragma unroll
      for (int k = 0; k < i; ++k) y[i] -= L[i][k] * y[k];
#pragma unroll
    for (int i = 0; i < 5; ++i) y[i] *= dinv[i];
#pragma unroll
    for (int i = 3; i >= 0; --i)
#pragma unroll
      for (int k = i + 1; k < 5; ++k) y[i] -= L[k][i] * y[k];
  };
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    double x[5], r[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) x[i] = double(S.b[f][i]);
    solve(x);
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      double t = double(S.b[f][i]);
#pragma unroll
      for (int c = 0; c < 5; ++c) t -= sym5(S.M, i, c) * x[c];
      r[i] = t;
    }
    solve(r);
#pragma unroll
    for (int i = 0; i < 5; ++i) X[f][i] = x[i] + r[i];
  }
  unscale_squares<NF>(X);
  return true;
}

// ---------------------------------------------------------------------------
// LS type A + stage update
// ---------------------------------------------------------------------------
struct LsaArgs {
  const Rec *__restrict__ mir;        // records in slot order (read)
  Rec *__restrict__ mout;             // records in slot order (write)
  const float4 *__restrict__ nbf;     // mixed path: FP32 neighbor fields of mir (else null)
  float4 *__restrict__ nbf_out;       // mixed path: those of mout (else null)
  const int32_t *__restrict__ sinfo;  // info in slot order
  const int32_t *__restrict__ items;  // slot -> physical id
  const int32_t *__restrict__ orig_of;  // physical id -> upload index
  const int32_t *__restrict__ cstart;
  const InstDev *__restrict__ inst;
  double *old_[6];
  double *new_[6];
  double *rhs[4];
  double *spd;  // non-null in the last RK stage: |u| of the new state per slot
  unsigned long long *stats64;
  int32_t *counters;
  int32_t *fb_list;
  int32_t *fb_count;
  int32_t *fb_done;
  MoveOut mv;
  int n, kt, list_len, mode;
  int staged;  // wave-staged candidate scan allowed
};

template <int NF>
__device__ __forceinline__ void scale_derivs(const double (&X)[NF][5], const InstDev &I, double (&d)[NF][5]) {
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    d[f][0] = X[f][0] * I.inv_h;
    d[f][1] = X[f][1] * I.inv_h;
    d[f][2] = X[f][2] * I.inv_h2;
    d[f][3] = X[f][3] * I.inv_h2;
    d[f][4] = X[f][4] * I.inv_h2;
  }
}

// FP64 assembly + DGESV-semantics solve; d[f] = (d_x, d_y, d_xx, d_yy, d_xy)
// (LSSPH_A_mod.f90:199-240).
// q = 5, 9 or 14 (SPH_MODEL 3-5).
template <int CLS, bool DELTA, int KT, int Q = 5>
__device__ __forceinline__ void derivs64(const LsaArgs &A, const InstDev &I, int lc, int s, const Rec &me,
                                         const NbList &L,
                                         double (&d)[Layout<CLS, DELTA>::NF][5]) {
  constexpr int NF = Layout<CLS, DELTA>::NF;
  const int kt = A.kt;
  LsSums<double, NF, Q> S;
  S.zero();
  walk_neighbors(A.mir, A.cstart, I, lc, s, me, L,
                 [&](const Rec &nb, double dx, double dy, int) {
                   const double r = sqrt(sq_dist(dx, dy));
                   const double w = I.w_norm * kshape_w<double, KT>(kt, r * I.inv_h);
                   double a[Q];
                   basis_q<Q>(dx, dy, I, a);
                   double df[NF];
                   field_diffs<CLS, DELTA>(nb, me, df);
                   S.add(nb.vol * w, a, df);
                 });
  double M[Q][Q], X[NF][Q];
  unpack(S, M, X);
  lu_solve<Q, NF>(M, X);
  double X5[NF][5];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int i = 0; i < 5; ++i) X5[f][i] = X[f][i];
  scale_derivs<NF>(X5, I, d);
}

// Mixed-precision assembly + FP64 LDL^T with refinement; false -> fallback.
// Offsets come from FP64 positions (x_j - x_i scaled by 1/h, then rounded);
// the neighbor's u, v, pre and volume are gathered as FP32 (nbf).
template <int CLS, bool DELTA, int KT>
__device__ __forceinline__ bool derivs32(const LsaArgs &A, const InstDev &I, int lc, int s, const Rec &me,
                                         const NbList &L,
                                         double (&d)[Layout<CLS, DELTA>::NF][5]) {
  constexpr int NF = Layout<CLS, DELTA>::NF;
  const int kt = A.kt;
  LsSums<float, NF> S;
  S.zero();
  const float4 own = nb_fields(me);
  walk_neighbors<true>(A.mir, A.cstart, I, lc, s, me, L,
                 [&](const Rec &nb, double dx, double dy, int t) {
                   const float xs = float(dx * I.inv_h), ys = float(dy * I.inv_h);
                   const float q = radius_s(xs, ys);
                   const float4 g = A.nbf[t];
                   // the common factor w_norm of M and b cancels in the solve
                   const float vw = g.w * kshape_w<float, KT, true>(kt, q);
                   float a[5];
                   basis5s(xs, ys, a);
                   float df[NF];
                   field_diffs32<CLS, DELTA>(g, own, nb, me, df);
                   S.add_reduced(vw, a, df);
                 });
  S.complete();
  double X[NF][5];
  if (!ldlt_solve_refine<NF>(S, X, I.cond_max)) return false;
  scale_derivs<NF>(X, I, d);
  return true;
}

// RHS of the governing equations (LSSPH_A_mod.f90:246-268) and the stage
// update (RK2_mod.f90:77-107, RK4_mod.f90:216-286) of physical particle p.
// Evaluated without FMA contraction, like the reference's arithmetic
// (baseline x86-64 has no FMA), so that every kernel that inlines it -- the
// operators and their FP64 redo -- rounds it identically.
template <int CLS, bool DELTA>
__device__ __forceinline__ void finish(const LsaArgs &A, const InstDev &I, int p, int s, const Rec &me,
                                       const double (&d)[Layout<CLS, DELTA>::NF][5]) {
#pragma clang fp contract(off)
  using L = Layout<CLS, DELTA>;
  double r_u, r_v = 0.0, r_rho = 0.0, r_tem = 0.0;
  if constexpr (CLS == 0) {
    r_u = I.diff_coef * (d[L::U][2] + d[L::U][3]) +
          I.diff_source * (2.0 * kPi * kPi * sin(kPi * (me.x - I.wl_thick)) * cos(kPi * (me.y - I.wl_thick)));
  } else {
    double drho_lap = 0.0;
    if constexpr (DELTA) drho_lap = d[L::RHO][2] + d[L::RHO][3];
    r_rho = -me.rho * (d[L::U][0] + d[L::V][1]) + I.delta * drho_lap;
    r_u = -d[L::P][0] / me.rho + (I.vis_ref / me.rho) * (d[L::U][2] + d[L::U][3]);
    r_v = -d[L::P][1] / me.rho + (I.vis_ref / me.rho) * (d[L::V][2] + d[L::V][3]);
    r_rho = r_rho / I.zeta_sq;
    if constexpr (CLS == 2) {
      r_v = r_v + I.alpha_ref * (me.tem - I.tem_ave) * I.gravity;
      r_u = r_u / I.xi_sq;
      r_v = r_v / I.xi_sq;
      r_tem = I.thermal_dif * (d[L::T][2] + d[L::T][3]);
    }
  }
  const int mode = A.mode;
  if (mode == kRhsOnly) {
    A.rhs[0][p] = r_u;
    A.rhs[1][p] = r_v;
    A.rhs[2][p] = r_rho;
    A.rhs[3][p] = r_tem;
    return;
  }
  constexpr bool DIFF = (CLS == 0);
  constexpr bool T = (CLS == 2);
  const bool per = I.periodic;
  const double dt = I.dt;
  Rec out = me;
  if (mode == kRk2S1 || mode == kRk2S2) {
    const double dte = (mode == kRk2S1) ? I.dt_half : dt;
    double ox, oy, ou, ov, orho, otem;
    if (mode == kRk2S1) {
      ox = me.x; oy = me.y; ou = me.u; ov = me.v; orho = me.rho; otem = me.tem;
      A.old_[2][p] = ou;
      if (!DIFF) {
        A.old_[0][p] = ox; A.old_[1][p] = oy; A.old_[3][p] = ov; A.old_[4][p] = orho;
      }
      if (T) A.old_[5][p] = otem;
    } else {
      ou = A.old_[2][p];
      ox = oy = ov = orho = otem = 0.0;
      if (!DIFF) {
        ox = A.old_[0][p]; oy = A.old_[1][p]; ov = A.old_[3][p]; orho = A.old_[4][p];
      }
      if (T) otem = A.old_[5][p];
    }
    if (!DIFF) {
      out.x = wrap_pos(ox + me.u * dte, I.len_x, per);
      out.y = wrap_pos(oy + me.v * dte, I.len_y, per);
      out.v = ov + r_v * dte;
      out.rho = orho + r_rho * dte;
    }
    out.u = ou + r_u * dte;
    if (T) out.tem = otem + r_tem * dte;
  } else {
    const int st = mode - kRk4S1 + 1;
    const double dt1 = (st == 1 || st == 4) ? I.dt_sixth : I.dt_third;
    const double dt2 = (st <= 2) ? I.dt_half : (st == 3 ? dt : 0.0);
    double nx_, ny_, nu_, nv_, nrho_, ntem_;
    if (st == 1) {
      A.old_[2][p] = me.u;
      if (!DIFF) {
        A.old_[0][p] = me.x; A.old_[1][p] = me.y; A.old_[3][p] = me.v; A.old_[4][p] = me.rho;
      }
      if (T) A.old_[5][p] = me.tem;
      nx_ = me.x; ny_ = me.y; nu_ = me.u; nv_ = me.v; nrho_ = me.rho; ntem_ = me.tem;
    } else {
      nu_ = A.new_[2][p];
      nx_ = ny_ = nv_ = nrho_ = ntem_ = 0.0;
      if (!DIFF) {
        nx_ = A.new_[0][p]; ny_ = A.new_[1][p]; nv_ = A.new_[3][p]; nrho_ = A.new_[4][p];
      }
      if (T) ntem_ = A.new_[5][p];
    }
    nu_ = nu_ + r_u * dt1;
    if (!DIFF) {
      nv_ = nv_ + r_v * dt1;
      nx_ = nx_ + me.u * dt1;
      ny_ = ny_ + me.v * dt1;
      nrho_ = nrho_ + r_rho * dt1;
    }
    if (T) ntem_ = ntem_ + r_tem * dt1;
    if (st < 4) {
      A.new_[2][p] = nu_;
      if (!DIFF) {
        A.new_[0][p] = nx_; A.new_[1][p] = ny_; A.new_[3][p] = nv_; A.new_[4][p] = nrho_;
      }
      if (T) A.new_[5][p] = ntem_;
      double ox, oy, ou, ov, orho, otem;
      if (st == 1) {
        ox = me.x; oy = me.y; ou = me.u; ov = me.v; orho = me.rho; otem = me.tem;
      } else {
        ou = A.old_[2][p];
        ox = oy = ov = orho = otem = 0.0;
        if (!DIFF) {
          ox = A.old_[0][p]; oy = A.old_[1][p]; ov = A.old_[3][p]; orho = A.old_[4][p];
        }
        if (T) otem = A.old_[5][p];
      }
      if (!DIFF) {
        out.x = wrap_pos(ox + me.u * dt2, I.len_x, per);
        out.y = wrap_pos(oy + me.v * dt2, I.len_y, per);
        out.v = ov + r_v * dt2;
        out.rho = orho + r_rho * dt2;
      }
      out.u = ou + r_u * dt2;
      if (T) out.tem = otem + r_tem * dt2;
    } else {
      out.u = nu_;
      if (!DIFF) {
        out.v = nv_;
        out.x = wrap_pos(nx_, I.len_x, per);
        out.y = wrap_pos(ny_, I.len_y, per);
        out.rho = nrho_;
      }
      if (T) out.tem = ntem_;
    }
  }
  if (!DIFF) {
    out.pre = I.c_sq * (out.rho - I.rho_ref);  // equation of state
    out.vol = I.sp_mass / out.rho;
    if (A.spd) A.spd[s] = sqrt(out.u * out.u + out.v * out.v);  // PST_mod.f90:51
  }
  store_rec(A.mout, A.nbf_out, s, out);
  if (!DIFF) record_move(A.mv, I, s, out.x, out.y);
}

// Arguments of the particle-shifting kernels.
struct PstArgs {
  const Rec *__restrict__ mir;
  Rec *__restrict__ mout;
  const float4 *__restrict__ nbf;     // mixed path: FP32 neighbor fields of mir (else null)
  float4 *__restrict__ nbf_out;       // mixed path: those of mout (else null)
  const int32_t *__restrict__ sinfo;
  const int32_t *__restrict__ items;
  const int32_t *__restrict__ orig_of;
  const int32_t *__restrict__ cstart;
  const InstDev *__restrict__ inst;
  const unsigned long long *__restrict__ umax;
  unsigned long long *stats64;
  int32_t *fb_list;   // slots queued for the FP64 redo (mixed path)
  int32_t *fb_count;
  int32_t *fb_done;
  MoveOut mv;
  int n, kt, list_len;
  int staged;  // wave-staged candidate scan allowed
};

// The LS operators of SPH_MODEL 6 (q = 20) on lane groups or row groups
// of a wavefront (highorder.hip).  cls: 0 diffusion, 1 flow, 2 Boussinesq.
void launch_lsa_wave(hipStream_t s, Device &d, const LsaArgs &a, int sph_model, int cls);
void launch_pst_wave(hipStream_t s, Device &d, const PstArgs &a, int sph_model);

}  // namespace lssph
