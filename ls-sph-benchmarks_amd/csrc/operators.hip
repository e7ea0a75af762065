// operators.hip -- LS-SPH pair operators for gfx950.
//
// One thread per mirror slot.  A pair kernel scans the stencil's candidates
// (x/y only, 16 B each) and compacts the accepted slots into a per-thread
// list in LDS (lds[k * blockDim + tid], bank-conflict free), then walks that
// list with every lane busy, loading 64-byte neighbor records; the ~100-flop
// pair body thus stays out of the 1-in-3 divergent candidate test.  A thread
// whose list overflows walks the candidates again directly.
//
// The LS type-A kernel (LSSPH_A_mod.f90:26-271) assembles the 5x5 moment
// matrix and right-hand sides in registers, solves, forms the RHS of the
// governing equations and applies the RK2/RK4 stage update (RK2_mod.f90:
// 83-110, RK4_mod.f90:234-288) in its epilogue, writing the particle's new
// record to the other physical buffer.
//   FP64 path: LU with partial pivoting (DGESV semantics).
//   Mixed path: FP32 pair sums on FP64-formed offsets and differences, FP64
//   LDL^T + one refinement step; particles whose LDL^T pivots are not
//   positive or whose pivot ratio exceeds cond_max are queued and redone by
//   the FP64 path in a second, list-driven launch.
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
// right-hand sides of the q = 5 LS fit.
template <typename R, int NF>
struct LsSums {
  R M[15];
  R b[NF][5];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < 15; ++i) M[i] = R(0);
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int i = 0; i < 5; ++i) b[f][i] = R(0);
  }
  // LSSPH_A_mod.f90:174-190
  __device__ __forceinline__ void add(R vw, const R (&a)[5], const R (&df)[NF]) {
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      const R vwa = vw * a[i];
#pragma unroll
      for (int f = 0; f < NF; ++f) b[f][i] += vwa * df[f];
#pragma unroll
      for (int r = 0; r <= i; ++r) M[i * (i + 1) / 2 + r] += vwa * a[r];
    }
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

// Same basis from offsets already scaled by 1/h (mixed precision).
__device__ __forceinline__ void basis5s(float xs, float ys, float (&a)[5]) {
  a[0] = xs;
  a[1] = ys;
  a[2] = 0.5f * (xs * xs);
  a[3] = 0.5f * (ys * ys);
  a[4] = xs * ys;
}

template <typename R, int NF>
__device__ __forceinline__ void unpack(const LsSums<R, NF> &S, double (&A)[5][5], double (&B)[NF][5]) {
#pragma unroll
  for (int i = 0; i < 5; ++i)
#pragma unroll
    for (int r = 0; r <= i; ++r) {
      A[r][i] = double(S.M[i * (i + 1) / 2 + r]);
      A[i][r] = A[r][i];
    }
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int i = 0; i < 5; ++i) B[f][i] = double(S.b[f][i]);
}

// Symmetric 5x5 entry (i, c) of packed upper-triangle sums.
template <typename R>
__device__ __forceinline__ double sym5(const R (&M)[15], int i, int c) {
  return i >= c ? double(M[i * (i + 1) / 2 + c]) : double(M[c * (c + 1) / 2 + i]);
}

// FP64 LDL^T of the symmetric 5x5 moment matrix held as FP32 sums, with one
// step of iterative refinement per right-hand side (x += M^-1 (b - M x),
// the residual formed in FP64 against the FP32 sums).  The right-hand sides
// are processed one at a time and the sums stay FP32 in registers, which
// keeps the solve's register footprint near that of the pair loop.  Returns
// false when a pivot is not positive or the pivot ratio max(d)/min(d) -- a
// cheap lower bound of cond_2(M) -- exceeds cond_max; the caller then
// re-solves in full FP64.
template <int NF>
__device__ __forceinline__ bool ldlt_solve_refine(const LsSums<float, NF> &S, double (&X)[NF][5],
                                                  double cond_max) {
  double L[5][5], d[5], dinv[5];
  double dmax = 0.0, dmin = 1.0e300;
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    double dj = sym5(S.M, j, j);
#pragma unroll
    for (int k = 0; k < j; ++k) dj -= L[j][k] * L[j][k] * d[k];
    ok = ok && (dj > 0.0);
    d[j] = dj;
    dinv[j] = 1.0 / dj;
    dmax = fmax(dmax, dj);
    dmin = fmin(dmin, dj);
#pragma unroll
    for (int i = j + 1; i < 5; ++i) {
      double t = sym5(S.M, i, j);
#pragma unroll
      for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k] * d[k];
      L[i][j] = t * dinv[j];
    }
  }
  if (!ok || dmax > cond_max * dmin) return false;
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
  return true;
}

// ---------------------------------------------------------------------------
// LS type A + stage update
// ---------------------------------------------------------------------------
struct LsaArgs {
  const Rec *__restrict__ mir;        // records in slot order (read)
  Rec *__restrict__ mout;             // records in slot order (write)
  const int32_t *__restrict__ sinfo;  // info in slot order
  const int32_t *__restrict__ items;  // slot -> physical id
  const int32_t *__restrict__ cstart;
  const InstDev *__restrict__ inst;
  double *old_[6];
  double *new_[6];
  double *rhs[4];
  unsigned long long *stats64;
  int32_t *counters;
  int32_t *fb_list;
  int32_t *fb_count;
  MoveOut mv;
  int n, kt, list_len, mode;
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
template <int CLS, bool DELTA, int KT>
__device__ __forceinline__ void derivs64(const LsaArgs &A, const InstDev &I, int lc, int s, const Rec &me,
                                         const int32_t *lds, int cnt, int list_len, int B, int tid,
                                         double (&d)[Layout<CLS, DELTA>::NF][5]) {
  constexpr int NF = Layout<CLS, DELTA>::NF;
  const int kt = A.kt;
  LsSums<double, NF> S;
  S.zero();
  walk_neighbors(A.mir, A.cstart, I, lc, s, me, lds, cnt, list_len, B, tid,
                 [&](const Rec &nb, double dx, double dy) {
                   const double r = sqrt(sq_dist(dx, dy));
                   const double w = I.w_norm * kshape_w<double, KT>(kt, r * I.inv_h);
                   double a[5];
                   basis5(dx, dy, I.inv_h, I.inv_h2, a);
                   double df[NF];
                   field_diffs<CLS, DELTA>(nb, me, df);
                   S.add(nb.vol * w, a, df);
                 });
  double M[5][5], X[NF][5];
  unpack(S, M, X);
  lu_solve<5, NF>(M, X);
  scale_derivs<NF>(X, I, d);
}

// Mixed-precision assembly + FP64 LDL^T with refinement; false -> fallback.
template <int CLS, bool DELTA, int KT>
__device__ __forceinline__ bool derivs32(const LsaArgs &A, const InstDev &I, int lc, int s, const Rec &me,
                                         const int32_t *lds, int cnt, int list_len, int B, int tid,
                                         double (&d)[Layout<CLS, DELTA>::NF][5]) {
  constexpr int NF = Layout<CLS, DELTA>::NF;
  const int kt = A.kt;
  LsSums<float, NF> S;
  S.zero();
  walk_neighbors(A.mir, A.cstart, I, lc, s, me, lds, cnt, list_len, B, tid,
                 [&](const Rec &nb, double dx, double dy) {
                   const float xs = float(dx * I.inv_h), ys = float(dy * I.inv_h);
                   const float q = fast_sqrtf(xs * xs + ys * ys);
                   const float vw = float(nb.vol * I.w_norm) * kshape_w<float, KT>(kt, q);
                   float a[5];
                   basis5s(xs, ys, a);
                   float df[NF];
                   field_diffs<CLS, DELTA>(nb, me, df);
                   S.add(vw, a, df);
                 });
  double X[NF][5];
  if (!ldlt_solve_refine<NF>(S, X, I.cond_max)) return false;
  scale_derivs<NF>(X, I, d);
  return true;
}

// RHS of the governing equations (LSSPH_A_mod.f90:246-268) and the stage
// update (RK2_mod.f90:77-107, RK4_mod.f90:216-286) of physical particle p.
template <int CLS, bool DELTA>
__device__ __forceinline__ void finish(const LsaArgs &A, const InstDev &I, int p, int s, const Rec &me,
                                       const double (&d)[Layout<CLS, DELTA>::NF][5]) {
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
    const double dte = (mode == kRk2S1) ? dt / 2.0 : dt;
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
    const double dt1 = (st == 1 || st == 4) ? dt / 6.0 : dt / 3.0;
    const double dt2 = (st <= 2) ? dt / 2.0 : (st == 3 ? dt : 0.0);
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
  }
  A.mout[s] = out;
  if (!DIFF) record_move(A.mv, I, s, out.x, out.y);
}

template <int CLS, bool MIXED, bool DELTA, int KT>
__global__ void __launch_bounds__(256) k_lsa(LsaArgs A) {
  constexpr int NF = Layout<CLS, DELTA>::NF;
  extern __shared__ int32_t lds[];
  const int tid = threadIdx.x, B = blockDim.x;
  const int s = xcd_block() * B + tid;
  if (s >= A.n) return;
  const int32_t inf = A.sinfo[s];
  if (info_ptype(inf) != 0) {
    if (A.mode != kRhsOnly) A.mout[s] = A.mir[s];  // walls move on with the records
    return;
  }
  const InstDev &I = A.inst[info_inst(inf)];
  const Rec me = A.mir[s];
  const int lc = local_cell(I, me.x, me.y);
  // list entries, the particle itself included
  const int cnt = scan_candidates(A.mir, A.cstart, I, lc, me.x, me.y, I.h_eff_sq, lds, A.list_len, B, tid);
  if (cnt > A.list_len) {
    atomicAdd(&A.stats64[kOverflows], 1ull);
    atomicMax(&A.counters[kMaxNbr], cnt - 1);
  }
  double d[NF][5];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int i = 0; i < 5; ++i) d[f][i] = 0.0;
  if (cnt - 1 >= 5) {  // LSSPH_A_mod.f90:208 (num_neighbors >= q)
    if constexpr (MIXED) {
      if (!derivs32<CLS, DELTA, KT>(A, I, lc, s, me, lds, cnt, A.list_len, B, tid, d)) {
        const int k = atomicAdd(A.fb_count, 1);
        A.fb_list[k] = s;
        return;
      }
    } else {
      derivs64<CLS, DELTA, KT>(A, I, lc, s, me, lds, cnt, A.list_len, B, tid, d);
    }
  }
  // Re-read the particle's own record for the epilogue instead of keeping it
  // live across the solve (the compiler barrier forbids reusing the first
  // load), which lowers the kernel's peak register count.
  __asm__ volatile("" ::: "memory");
  const Rec me2 = A.mir[s];
  finish<CLS, DELTA>(A, I, A.items[s], s, me2, d);
}

// FP64 redo of the slots the mixed kernel queued (direct neighbor walk).
template <int CLS, bool DELTA, int KT>
__global__ void __launch_bounds__(256) k_lsa_fallback(LsaArgs A) {
  constexpr int NF = Layout<CLS, DELTA>::NF;
  const int count = *A.fb_count;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const int s = A.fb_list[i];
    const InstDev &I = A.inst[info_inst(A.sinfo[s])];
    const Rec me = A.mir[s];
    const int lc = local_cell(I, me.x, me.y);
    double d[NF][5];
    derivs64<CLS, DELTA, KT>(A, I, lc, s, me, nullptr, 1 << 30, 0, 0, 0, d);
    atomicAdd(&A.stats64[kFallbacks], 1ull);
    finish<CLS, DELTA>(A, I, A.items[s], s, me, d);
  }
}

// ---------------------------------------------------------------------------
// Particle shifting (PST_mod.f90:37-114, shift_interpolate_mod.f90:23-300)
// ---------------------------------------------------------------------------
struct PstArgs {
  const Rec *__restrict__ mir;
  Rec *__restrict__ mout;
  const int32_t *__restrict__ sinfo;
  const int32_t *__restrict__ items;
  const int32_t *__restrict__ cstart;
  const InstDev *__restrict__ inst;
  const unsigned long long *__restrict__ umax;
  unsigned long long *stats64;
  MoveOut mv;
  int n, kt, list_len;
};

template <bool T, bool MIXED, int KT>
__global__ void __launch_bounds__(256) k_pst(PstArgs A) {
  constexpr int NF = T ? 4 : 3;  // u, v, rho (, tem)
  using R = typename std::conditional<MIXED, float, double>::type;
  extern __shared__ int32_t lds[];
  const int tid = threadIdx.x, B = blockDim.x;
  const int s = xcd_block() * B + tid;
  if (s >= A.n) return;
  const int32_t inf = A.sinfo[s];
  if (info_ptype(inf) != 0) {
    A.mout[s] = A.mir[s];  // walls move on with the records
    return;
  }
  const int ii = info_inst(inf);
  const InstDev &I = A.inst[ii];
  const Rec me = A.mir[s];
  const int lc = local_cell(I, me.x, me.y);
  const int kt = A.kt;
  // list entries, the particle itself included
  const int cnt = scan_candidates(A.mir, A.cstart, I, lc, me.x, me.y, I.h_eff_sq, lds, A.list_len, B, tid);
  if (cnt > A.list_len) atomicAdd(&A.stats64[kOverflows], 1ull);

  LsSums<R, NF> S;
  S.zero();
  R sx = R(0), sy = R(0);
  const double inv_w_ave = 1.0 / I.w_ave;
  walk_neighbors(A.mir, A.cstart, I, lc, s, me, lds, cnt, A.list_len, B, tid,
                 [&](const Rec &nb, double dx, double dy) {
                   R df[NF];
                   df[0] = R(nb.u - me.u);
                   df[1] = R(nb.v - me.v);
                   df[2] = R(nb.rho - me.rho);
                   if constexpr (T) df[3] = R(nb.tem - me.tem);
                   if constexpr (!MIXED) {
                     const double r = sqrt(sq_dist(dx, dy));
                     const double q = r * I.inv_h;
                     const double w = I.w_norm * kshape_w<double, KT>(kt, q);
                     const double dwdr = I.w_norm * kshape_dw<double, KT>(kt, q) * I.dw_fac;
                     const double vj = nb.vol;
                     const double inv_r = 1.0 / r;
                     // the particle itself (r = 0) adds no shift
                     const double dwx = r > 0.0 ? dwdr * (-dx) * inv_r * vj : 0.0;
                     const double dwy = r > 0.0 ? dwdr * (-dy) * inv_r * vj : 0.0;
                     const double wr = w * inv_w_ave;
                     const double f = 1.0 + 0.2 * ((wr * wr) * (wr * wr));
                     sx = sx + f * dwx;
                     sy = sy + f * dwy;
                     double a[5];
                     basis5(dx, dy, I.inv_h, I.inv_h2, a);
                     S.add(vj * w, a, df);
                   } else {
                     const float xs = float(dx * I.inv_h), ys = float(dy * I.inv_h);
                     const float q = fast_sqrtf(xs * xs + ys * ys);
                     const float shp = kshape_w<float, KT>(kt, q);
                     const float vwn = float(nb.vol * I.w_norm);
                     // dW/dr * (-x/r) * V = V * w_norm * dw_fac * shape'(q) * (-xs / q)
                     const float g0 = float(nb.vol * I.w_norm * I.dw_fac) * kshape_dw<float, KT>(kt, q) * fast_rcpf(q);
                     const float g = q > 0.0f ? g0 : 0.0f;  // the particle itself adds no shift
                     const float wr = float(I.w_norm * inv_w_ave) * shp;
                     const float f = 1.0f + 0.2f * ((wr * wr) * (wr * wr));
                     sx = sx + f * (g * -xs);
                     sy = sy + f * (g * -ys);
                     float a[5];
                     basis5s(xs, ys, a);
                     S.add(vwn * shp, a, df);
                   }
                 });
  // shift vector and its clamp (shift_interpolate_mod.f90:333-346)
  const double umax = __longlong_as_double((long long)A.umax[ii]);
  const double coe = I.pst_c * umax * I.dt * I.h;
  double dxs = -coe * double(sx), dys = -coe * double(sy);
  const double rs = sqrt(dxs * dxs + dys * dys);
  if (rs > 0.2 * I.h) {
    dxs = (0.2 * I.h) * dxs / rs;
    dys = (0.2 * I.h) * dys / rs;
  }
  // LS correction of u, v, rho (, tem) to the shifted position (:352-415)
  double X[NF][5];
  if (cnt - 1 >= 5) {  // shift_interpolate_mod.f90:240 (num_neighbors >= q)
    bool solved = false;
    if constexpr (MIXED) {
      solved = ldlt_solve_refine<NF>(S, X, I.cond_max);
      if (!solved) atomicAdd(&A.stats64[kFallbacks], 1ull);
    }
    if (!solved) {
      double M[5][5];
      unpack(S, M, X);
      lu_solve<5, NF>(M, X);
    }
  } else {
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int i = 0; i < 5; ++i) X[f][i] = 0.0;
  }
  double a[5];
  basis5(dxs, dys, I.inv_h, I.inv_h2, a);
  double acc[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) t += a[i] * X[f][i];
    acc[f] = t;
  }
  Rec out = me;
  const bool per = I.periodic;
  out.x = wrap_pos(me.x + dxs, I.len_x, per);
  out.y = wrap_pos(me.y + dys, I.len_y, per);
  out.u = me.u + acc[0];
  out.v = me.v + acc[1];
  out.rho = me.rho + acc[2];
  if constexpr (T) out.tem = me.tem + acc[3];
  out.pre = I.c_sq * (out.rho - I.rho_ref);
  out.vol = I.sp_mass / out.rho;
  A.mout[s] = out;
  record_move(A.mv, I, s, out.x, out.y);
}

// ---------------------------------------------------------------------------
// Per-instance max reductions over the fluid particles of a physical range
// (an instance occupies the contiguous physical range [pbeg, pend)).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_max(double v) {
  __shared__ double sh[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double m = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) m = fmax(m, sh[i]);
  return m;
}

// U_max of PST_mod.f90:49-53.
// An instance's particles occupy the contiguous slot range of its cells.
__global__ void __launch_bounds__(256) k_umax(const Rec *__restrict__ rec, const int32_t *__restrict__ sinfo,
                                              const int32_t *__restrict__ cstart, const InstDev *__restrict__ inst,
                                              unsigned long long *umax) {
  const InstDev &I = inst[blockIdx.y];
  const int t0 = cstart[I.cell_off], t1 = cstart[I.cell_off + I.nx * I.ny];
  double s = 0.0;
  for (int t = t0 + blockIdx.x * blockDim.x + threadIdx.x; t < t1; t += gridDim.x * blockDim.x) {
    if (info_ptype(sinfo[t]) == 0) {
      const double2 uv = *reinterpret_cast<const double2 *>(&rec[t].u);
      s = fmax(s, sqrt(uv.x * uv.x + uv.y * uv.y));
    }
  }
  const double m = block_max(s);
  if (threadIdx.x == 0) atomic_max_nonneg(&umax[blockIdx.y], m);
}

// max|f_previous - u| of check_steady_state_mod.f90:36-46 (f_previous in
// physical order).
__global__ void __launch_bounds__(256) k_dfmax(const Rec *__restrict__ rec, const int32_t *__restrict__ sinfo,
                                               const int32_t *__restrict__ items,
                                               const int32_t *__restrict__ cstart, const double *__restrict__ f_prev,
                                               const InstDev *__restrict__ inst, unsigned long long *dfmax) {
  const InstDev &I = inst[blockIdx.y];
  const int t0 = cstart[I.cell_off], t1 = cstart[I.cell_off + I.nx * I.ny];
  double s = 0.0;
  for (int t = t0 + blockIdx.x * blockDim.x + threadIdx.x; t < t1; t += gridDim.x * blockDim.x)
    if (info_ptype(sinfo[t]) == 0) s = fmax(s, fabs(f_prev[items[t]] - rec[t].u));
  const double m = block_max(s);
  if (threadIdx.x == 0) atomic_max_nonneg(&dfmax[blockIdx.y], m);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
// LDS of the neighbor lists: list_len 32-bit entries per thread plus the
// trash entry of the scan.
static size_t list_lds_bytes(int list_len, int block) { return (size_t)4 * (list_len + 1) * block; }

template <int CLS, bool MIXED, bool DELTA, int KT>
static void lsa_launch(hipStream_t s, const LsaArgs &a, int block, size_t lds) {
  hipLaunchKernelGGL((k_lsa<CLS, MIXED, DELTA, KT>), dim3(xcd_grid(a.n, block)), dim3(block), lds, s, a);
  if (MIXED) hipLaunchKernelGGL((k_lsa_fallback<CLS, DELTA, KT>), dim3(64), dim3(128), 0, s, a);
}

// Wendland C2 (the default kernel) is compiled in; other kernels take the
// run-time switch.
template <int CLS, bool DELTA>
static void lsa_dispatch(hipStream_t s, const LsaArgs &a, int block, size_t lds, bool mixed, int kt) {
  if (kt == 3) {
    if (mixed)
      lsa_launch<CLS, true, DELTA, 3>(s, a, block, lds);
    else
      lsa_launch<CLS, false, DELTA, 3>(s, a, block, lds);
  } else {
    if (mixed)
      lsa_launch<CLS, true, DELTA, 0>(s, a, block, lds);
    else
      lsa_launch<CLS, false, DELTA, 0>(s, a, block, lds);
  }
}

void launch_lsa(hipStream_t s, Device &d, int cur, StageMode mode, int guard, int keys) {
  LsaArgs a;
  a.mir = d.mir[cur];
  a.mout = d.mir[cur ^ 1];
  a.sinfo = d.sinfo;
  a.items = d.items;
  a.cstart = d.cell_start;
  a.inst = d.inst;
  for (int f = 0; f < 6; ++f) {
    a.old_[f] = d.old_[f];
    a.new_[f] = d.new_[f];
  }
  for (int f = 0; f < 4; ++f) a.rhs[f] = d.rhs[f];
  a.stats64 = d.stats64;
  a.counters = d.counters;
  a.fb_list = d.fb_list;
  a.fb_count = d.counters + kFbCount;
  a.mv = MoveOut{d.kslot, d.strays, d.nstray, d.counters, d.xbuild, guard, keys};
  a.n = d.n_total;
  a.kt = d.opt.kernel_type;
  a.list_len = d.list_len;
  a.mode = mode;
  const size_t lds = list_lds_bytes(d.list_len, d.block);
  const bool mixed = d.opt.precision == LSSPH_MIXED;
  if (mixed) launch_fill_i32(s, a.fb_count, 0, 1);
  const int cls = d.opt.target_problem == 1 ? 0 : (d.opt.has_tem ? 2 : 1);
  const bool delta = d.opt.delta != 0;
  const int kt = d.opt.kernel_type;
  if (cls == 0)
    lsa_dispatch<0, false>(s, a, d.block, lds, mixed, kt);
  else if (cls == 1)
    delta ? lsa_dispatch<1, true>(s, a, d.block, lds, mixed, kt) : lsa_dispatch<1, false>(s, a, d.block, lds, mixed, kt);
  else
    delta ? lsa_dispatch<2, true>(s, a, d.block, lds, mixed, kt) : lsa_dispatch<2, false>(s, a, d.block, lds, mixed, kt);
}

void launch_umax(hipStream_t s, Device &d, int cur) {
  launch_fill_u64(s, d.umax, 0ull, d.n_inst);
  hipLaunchKernelGGL(k_umax, dim3(d.red_blocks, d.n_inst), dim3(256), 0, s, d.mir[cur], d.sinfo, d.cell_start,
                     d.inst, d.umax);
}

void launch_dfmax(hipStream_t s, Device &d, int cur) {
  launch_fill_u64(s, d.dfmax, 0ull, d.n_inst);
  hipLaunchKernelGGL(k_dfmax, dim3(d.red_blocks, d.n_inst), dim3(256), 0, s, d.mir[cur], d.sinfo, d.items,
                     d.cell_start, d.f_prev, d.inst, d.dfmax);
}

template <bool T, bool MIXED>
static void pst_launch(hipStream_t s, const PstArgs &a, int block, size_t lds, int kt) {
  const dim3 g(xcd_grid(a.n, block)), b(block);
  if (kt == 3)
    hipLaunchKernelGGL((k_pst<T, MIXED, 3>), g, b, lds, s, a);
  else
    hipLaunchKernelGGL((k_pst<T, MIXED, 0>), g, b, lds, s, a);
}

void launch_pst(hipStream_t s, Device &d, int cur, int guard, int keys) {
  PstArgs a;
  a.mir = d.mir[cur];
  a.mout = d.mir[cur ^ 1];
  a.sinfo = d.sinfo;
  a.items = d.items;
  a.cstart = d.cell_start;
  a.inst = d.inst;
  a.umax = d.umax;
  a.stats64 = d.stats64;
  a.mv = MoveOut{d.kslot, d.strays, d.nstray, d.counters, d.xbuild, guard, keys};
  a.n = d.n_total;
  a.kt = d.opt.kernel_type;
  a.list_len = d.list_len;
  const size_t lds = list_lds_bytes(d.list_len, d.block);
  const bool mixed = d.opt.precision == LSSPH_MIXED;
  const int kt = d.opt.kernel_type;
  if (d.opt.has_tem)
    mixed ? pst_launch<true, true>(s, a, d.block, lds, kt) : pst_launch<true, false>(s, a, d.block, lds, kt);
  else
    mixed ? pst_launch<false, true>(s, a, d.block, lds, kt) : pst_launch<false, false>(s, a, d.block, lds, kt);
}

}  // namespace lssph
