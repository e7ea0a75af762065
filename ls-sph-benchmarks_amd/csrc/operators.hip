// operators.hip -- LS-SPH pair operators for gfx950.
//
// One thread per particle (64-wide wavefronts).  A pair kernel first scans
// the candidates of the 3x3 cell stencil (x/y only, 16 B per candidate) and
// compacts the accepted neighbor ids into a per-thread list in LDS (layout
// lds[k * blockDim + tid], bank-conflict free), then walks that list with
// every lane busy, loading the 64-byte neighbor records.  This keeps the
// ~100-flop pair body out of the 1-in-3 divergent candidate test.  A thread
// whose list overflows redoes the walk directly over the candidates.
//
// The LS type-A kernel (LSSPH_A_mod.f90:26-271) assembles the 5x5 moment
// matrix and the right-hand sides in registers, solves with an unrolled LU
// (DGESV semantics), forms the RHS of the governing equations and applies
// the RK2/RK4 stage update (RK2_mod.f90:83-110, RK4_mod.f90:234-288) in its
// epilogue, writing the particle's new record to the other buffer.
#include <type_traits>

#include "internal.hpp"

namespace lssph {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
// x_j - x_i with the periodic image chosen as the oracle's wrapped cell does
// ((x_j + shift) - x_i); identical to the plain difference in a closed box.
__device__ __forceinline__ double pdelta(double xj, double xi, double len, bool periodic) {
  double d = xj - xi;
  if (periodic) {
    if (d > 0.5 * len)
      d = (xj - len) - xi;
    else if (d < -0.5 * len)
      d = (xj + len) - xi;
  }
  return d;
}

// r^2 without FMA contraction, as the reference forms it (LSSPH_A_mod.f90:
// 123), so the acceptance test r^2 <= h_eff^2 decides exactly as it does.
__device__ __forceinline__ double sq_dist(double dx, double dy) {
#pragma clang fp contract(off)
  return dx * dx + dy * dy;
}

__device__ __forceinline__ double wrap_pos(double x, double len, bool periodic) {
  if (!periodic) return x;
  if (x >= len) return x - len;
  if (x < 0.0) return x + len;
  return x;
}

__device__ __forceinline__ double2 load_xy(const Rec *__restrict__ r, int j) {
  return *reinterpret_cast<const double2 *>(&r[j].x);
}

// Visits the candidate ranges of the 3x3 stencil around local cell `lc` in
// the reference's order (rows dy = -1..1, within a row dx = -1..1).  Cells
// of one row are consecutive in items[], so a row is one range (two when a
// periodic row wraps).
template <typename F>
__device__ __forceinline__ void visit_cells(const InstDev &I, int lc, const int32_t *__restrict__ cstart,
                                            F &&f) {
  const int nx = I.nx, ny = I.ny;
  const int cx = lc % nx, cy = lc / nx;
#pragma unroll
  for (int dy = -1; dy <= 1; ++dy) {
    int yy = cy + dy;
    if (I.periodic) {
      yy = yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy);
    } else if (yy < 0 || yy >= ny) {
      continue;
    }
    const int row = I.cell_off + yy * nx;
    if (!I.periodic) {
      const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx < nx - 1 ? cx + 1 : nx - 1;
      f(cstart[row + x0], cstart[row + x1 + 1]);
    } else if (cx == 0) {
      f(cstart[row + nx - 1], cstart[row + nx]);
      f(cstart[row], cstart[row + 2]);
    } else if (cx == nx - 1) {
      f(cstart[row + cx - 1], cstart[row + cx + 1]);
      f(cstart[row], cstart[row + 1]);
    } else {
      f(cstart[row + cx - 1], cstart[row + cx + 2]);
    }
  }
}

// Fields carried by the LS type-A sums: diffusion -> u only; flow -> u, v,
// rho, pre; Boussinesq -> u, v, rho, pre, tem.
template <int CLS> struct Fields { static constexpr int NF = CLS == 0 ? 1 : (CLS == 1 ? 4 : 5); };

template <int NF, typename R>
__device__ __forceinline__ void field_diffs(const Rec &nb, const Rec &me, R (&df)[NF]) {
  df[0] = R(nb.u - me.u);
  if constexpr (NF > 1) {
    df[1] = R(nb.v - me.v);
    df[2] = R(nb.rho - me.rho);
    df[3] = R(nb.pre - me.pre);
  }
  if constexpr (NF > 4) df[4] = R(nb.tem - me.tem);
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

// FP64 pair contribution.
template <int NF>
__device__ __forceinline__ void add_pair64(LsSums<double, NF> &S, const InstDev &I, int kt, const Rec &me,
                                           const Rec &nb, double dx, double dy) {
  const double r = sqrt(sq_dist(dx, dy));
  const double w = I.w_norm * kshape_w<double>(kt, r * I.inv_h);
  const double vw = nb.vol * w;
  double a[5];
  basis5(dx, dy, I.inv_h, I.inv_h2, a);
  double df[NF];
  field_diffs<NF>(nb, me, df);
  S.add(vw, a, df);
}

// Mixed-precision pair contribution: offsets and field differences formed in
// FP64, scaled by 1/h, then rounded to FP32 for the products and sums.
template <int NF>
__device__ __forceinline__ void add_pair32(LsSums<float, NF> &S, const InstDev &I, int kt, const Rec &me,
                                           const Rec &nb, double dx, double dy) {
  const float xs = float(dx * I.inv_h), ys = float(dy * I.inv_h);
  const float q = sqrtf(xs * xs + ys * ys);
  const float vw = float(nb.vol * I.w_norm) * kshape_w<float>(kt, q);
  float a[5];
  basis5s(xs, ys, a);
  float df[NF];
  field_diffs<NF>(nb, me, df);
  S.add(vw, a, df);
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

// Walks the neighbors of particle p (LDS list, or the candidates directly on
// overflow) and calls f(nb_record, dx, dy) for each accepted neighbor in
// summation order.
template <typename F>
__device__ __forceinline__ void walk_neighbors(const Rec *__restrict__ rin, const int32_t *__restrict__ items,
                                               const int32_t *__restrict__ cstart, const InstDev &I, int lc,
                                               int p, const Rec &me, const int32_t *lds, int cnt, int list_len,
                                               int B, int tid, F &&f) {
  const bool per = I.periodic;
  if (cnt <= list_len) {
    for (int k = 0; k < cnt; ++k) {
      const int j = lds[k * B + tid];
      const Rec nb = rin[j];
      f(nb, pdelta(nb.x, me.x, I.len_x, per), pdelta(nb.y, me.y, I.len_y, per));
    }
  } else {
    visit_cells(I, lc, cstart, [&](int t0, int t1) {
      for (int t = t0; t < t1; ++t) {
        const int j = items[t];
        if (j == p) continue;
        const double2 xy = load_xy(rin, j);
        const double dx = pdelta(xy.x, me.x, I.len_x, per), dy = pdelta(xy.y, me.y, I.len_y, per);
        if (sq_dist(dx, dy) <= I.h_eff_sq) f(rin[j], dx, dy);
      }
    });
  }
}

// Candidate scan: fills the LDS list, returns the neighbor count.
__device__ __forceinline__ int scan_candidates(const Rec *__restrict__ rin, const int32_t *__restrict__ items,
                                               const int32_t *__restrict__ cstart, const InstDev &I, int lc,
                                               int p, double xi, double yi, int32_t *lds, int list_len, int B,
                                               int tid, const int32_t *__restrict__ info, bool fluid_only) {
  const bool per = I.periodic;
  int cnt = 0;
  visit_cells(I, lc, cstart, [&](int t0, int t1) {
    for (int t = t0; t < t1; ++t) {
      const int j = items[t];
      if (j == p) continue;
      if (fluid_only && info_ptype(info[j]) != 0) continue;
      const double2 xy = load_xy(rin, j);
      const double dx = pdelta(xy.x, xi, I.len_x, per), dy = pdelta(xy.y, yi, I.len_y, per);
      if (sq_dist(dx, dy) <= I.h_eff_sq) {
        if (cnt < list_len) lds[cnt * B + tid] = j;
        ++cnt;
      }
    }
  });
  return cnt;
}

// ---------------------------------------------------------------------------
// LS type A + stage update
// ---------------------------------------------------------------------------
struct LsaArgs {
  const Rec *__restrict__ rin;
  Rec *__restrict__ rout;
  const int32_t *__restrict__ info;
  const int32_t *__restrict__ key;
  const int32_t *__restrict__ cstart;
  const int32_t *__restrict__ items;
  const InstDev *__restrict__ inst;
  double *old_[6];
  double *new_[6];
  double *rhs[4];
  unsigned long long *stats64;
  int32_t *counters;
  int n, kt, list_len, mode;
};

// Solves the 5x5 system in FP64; returns the scaled derivatives d[f][0..4] =
// (d_x, d_y, d_xx, d_yy, d_xy) (LSSPH_A_mod.f90:203-228).
template <int NF>
__device__ __forceinline__ void derivs_from(double (&A)[5][5], double (&B)[NF][5], const InstDev &I,
                                            double (&d)[NF][5]) {
  lu_solve<5, NF>(A, B);
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    d[f][0] = B[f][0] * I.inv_h;
    d[f][1] = B[f][1] * I.inv_h;
    d[f][2] = B[f][2] * I.inv_h2;
    d[f][3] = B[f][3] * I.inv_h2;
    d[f][4] = B[f][4] * I.inv_h2;
  }
}

template <int CLS, bool MIXED>
__global__ void __launch_bounds__(256) k_lsa(LsaArgs A) {
  constexpr int NF = Fields<CLS>::NF;
  extern __shared__ int32_t lds[];
  const int tid = threadIdx.x, B = blockDim.x;
  const int p = blockIdx.x * B + tid;
  if (p >= A.n) return;
  const int32_t inf = A.info[p];
  if (info_ptype(inf) != 0) return;
  const InstDev &I = A.inst[info_inst(inf)];
  const Rec me = A.rin[p];
  const int lc = A.key[p] - I.cell_off;
  const int kt = A.kt;

  const int cnt = scan_candidates(A.rin, A.items, A.cstart, I, lc, p, me.x, me.y, lds, A.list_len, B, tid,
                                  A.info, false);
  if (cnt > A.list_len) {
    atomicAdd(&A.stats64[kOverflows], 1ull);
    atomicMax(&A.counters[kMaxNbr], cnt);
  }

  double d[NF][5];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int i = 0; i < 5; ++i) d[f][i] = 0.0;

  if (cnt >= 5) {  // LSSPH_A_mod.f90:208 (num_neighbors >= q)
    bool need64 = !MIXED;
    if constexpr (MIXED) {
      LsSums<float, NF> S;
      S.zero();
      walk_neighbors(A.rin, A.items, A.cstart, I, lc, p, me, lds, cnt, A.list_len, B, tid,
                     [&](const Rec &nb, double dx, double dy) { add_pair32<NF>(S, I, kt, me, nb, dx, dy); });
      double M[5][5], X[NF][5];
      unpack(S, M, X);
      int piv[5];
      double LU[5][5];
#pragma unroll
      for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int c = 0; c < 5; ++c) LU[i][c] = M[i][c];
      const int info = lu_factor<5, NF>(LU, piv);
      double umax = 0.0, umin = 1.0e300;
#pragma unroll
      for (int i = 0; i < 5; ++i) {
        const double t = fabs(LU[i][i]);
        umax = fmax(umax, t);
        umin = fmin(umin, t);
      }
      if (info != 0 || umax > I.cond_max * umin) {
        need64 = true;
        atomicAdd(&A.stats64[kFallbacks], 1ull);
      } else {
        double Bv[NF][5];
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
          for (int i = 0; i < 5; ++i) Bv[f][i] = X[f][i];
        lu_apply<5, NF>(LU, piv, X);
        // one step of iterative refinement: r = b - M x, x += M^-1 r
        double R[NF][5];
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
          for (int i = 0; i < 5; ++i) {
            double s = Bv[f][i];
#pragma unroll
            for (int c = 0; c < 5; ++c) s -= M[i][c] * X[f][c];
            R[f][i] = s;
          }
        lu_apply<5, NF>(LU, piv, R);
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          d[f][0] = (X[f][0] + R[f][0]) * I.inv_h;
          d[f][1] = (X[f][1] + R[f][1]) * I.inv_h;
          d[f][2] = (X[f][2] + R[f][2]) * I.inv_h2;
          d[f][3] = (X[f][3] + R[f][3]) * I.inv_h2;
          d[f][4] = (X[f][4] + R[f][4]) * I.inv_h2;
        }
      }
    }
    if (need64) {
      LsSums<double, NF> S;
      S.zero();
      walk_neighbors(A.rin, A.items, A.cstart, I, lc, p, me, lds, cnt, A.list_len, B, tid,
                     [&](const Rec &nb, double dx, double dy) { add_pair64<NF>(S, I, kt, me, nb, dx, dy); });
      double M[5][5], X[NF][5];
      unpack(S, M, X);
      derivs_from<NF>(M, X, I, d);
    }
  }

  // RHS of the governing equations, LSSPH_A_mod.f90:246-268
  double r_u, r_v = 0.0, r_rho = 0.0, r_tem = 0.0;
  if constexpr (CLS == 0) {
    r_u = I.diff_coef * (d[0][2] + d[0][3]) +
          I.diff_source * (2.0 * kPi * kPi * sin(kPi * (me.x - I.wl_thick)) * cos(kPi * (me.y - I.wl_thick)));
  } else {
    r_rho = -me.rho * (d[0][0] + d[1][1]) + I.delta * (d[2][2] + d[2][3]);
    r_u = -d[3][0] / me.rho + (I.vis_ref / me.rho) * (d[0][2] + d[0][3]);
    r_v = -d[3][1] / me.rho + (I.vis_ref / me.rho) * (d[1][2] + d[1][3]);
    r_rho = r_rho / I.zeta_sq;
    if constexpr (CLS == 2) {
      r_v = r_v + I.alpha_ref * (me.tem - I.tem_ave) * I.gravity;
      r_u = r_u / I.xi_sq;
      r_v = r_v / I.xi_sq;
      r_tem = I.thermal_dif * (d[4][2] + d[4][3]);
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

  // Stage update (RK2_mod.f90:77-107, RK4_mod.f90:216-286)
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
      A.old_[0][p] = ox; A.old_[1][p] = oy; A.old_[2][p] = ou;
      A.old_[3][p] = ov; A.old_[4][p] = orho;
      if (T) A.old_[5][p] = otem;
    } else {
      ox = A.old_[0][p]; oy = A.old_[1][p]; ou = A.old_[2][p];
      ov = A.old_[3][p]; orho = A.old_[4][p]; otem = T ? A.old_[5][p] : 0.0;
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
      A.old_[0][p] = me.x; A.old_[1][p] = me.y; A.old_[2][p] = me.u;
      A.old_[3][p] = me.v; A.old_[4][p] = me.rho;
      if (T) A.old_[5][p] = me.tem;
      nx_ = me.x; ny_ = me.y; nu_ = me.u; nv_ = me.v; nrho_ = me.rho; ntem_ = me.tem;
    } else {
      nx_ = A.new_[0][p]; ny_ = A.new_[1][p]; nu_ = A.new_[2][p];
      nv_ = A.new_[3][p]; nrho_ = A.new_[4][p]; ntem_ = T ? A.new_[5][p] : 0.0;
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
        ox = A.old_[0][p]; oy = A.old_[1][p]; ou = A.old_[2][p];
        ov = A.old_[3][p]; orho = A.old_[4][p]; otem = T ? A.old_[5][p] : 0.0;
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
  A.rout[p] = out;
}

// ---------------------------------------------------------------------------
// Particle shifting (PST_mod.f90:37-114, shift_interpolate_mod.f90:23-300)
// ---------------------------------------------------------------------------
struct PstArgs {
  const Rec *__restrict__ rin;
  Rec *__restrict__ rout;
  const int32_t *__restrict__ info;
  const int32_t *__restrict__ key;
  const int32_t *__restrict__ cstart;
  const int32_t *__restrict__ items;
  const InstDev *__restrict__ inst;
  const unsigned long long *__restrict__ umax;
  unsigned long long *stats64;
  int n, kt, list_len, has_tem, mixed;
};

template <bool T, bool MIXED>
__global__ void __launch_bounds__(256) k_pst(PstArgs A) {
  constexpr int NF = T ? 4 : 3;  // u, v, rho (, tem)
  using R = typename std::conditional<MIXED, float, double>::type;
  extern __shared__ int32_t lds[];
  const int tid = threadIdx.x, B = blockDim.x;
  const int p = blockIdx.x * B + tid;
  if (p >= A.n) return;
  const int32_t inf = A.info[p];
  if (info_ptype(inf) != 0) return;
  const int ii = info_inst(inf);
  const InstDev &I = A.inst[ii];
  const Rec me = A.rin[p];
  const int lc = A.key[p] - I.cell_off;
  const int kt = A.kt;
  const int cnt = scan_candidates(A.rin, A.items, A.cstart, I, lc, p, me.x, me.y, lds, A.list_len, B, tid,
                                  A.info, false);
  if (cnt > A.list_len) atomicAdd(&A.stats64[kOverflows], 1ull);

  LsSums<R, NF> S;
  S.zero();
  R sx = R(0), sy = R(0);
  const double inv_w_ave = 1.0 / I.w_ave;
  walk_neighbors(A.rin, A.items, A.cstart, I, lc, p, me, lds, cnt, A.list_len, B, tid,
                 [&](const Rec &nb, double dx, double dy) {
                   R df[NF];
                   df[0] = R(nb.u - me.u);
                   df[1] = R(nb.v - me.v);
                   df[2] = R(nb.rho - me.rho);
                   if constexpr (T) df[3] = R(nb.tem - me.tem);
                   if constexpr (!MIXED) {
                     const double r = sqrt(sq_dist(dx, dy));
                     const double q = r * I.inv_h;
                     const double w = I.w_norm * kshape_w<double>(kt, q);
                     const double dwdr = I.w_norm * kshape_dw<double>(kt, q) * I.dw_fac;
                     const double vj = nb.vol;
                     const double inv_r = 1.0 / r;
                     const double dwx = dwdr * (-dx) * inv_r * vj;
                     const double dwy = dwdr * (-dy) * inv_r * vj;
                     const double wr = w * inv_w_ave;
                     const double f = 1.0 + 0.2 * ((wr * wr) * (wr * wr));
                     sx = sx + f * dwx;
                     sy = sy + f * dwy;
                     double a[5];
                     basis5(dx, dy, I.inv_h, I.inv_h2, a);
                     S.add(vj * w, a, df);
                   } else {
                     const float xs = float(dx * I.inv_h), ys = float(dy * I.inv_h);
                     const float q = sqrtf(xs * xs + ys * ys);
                     const float shp = kshape_w<float>(kt, q);
                     const float vwn = float(nb.vol * I.w_norm);
                     // dW/dr * (-x/r) = w_norm * shape'(q) * dw_fac * (-xs/q)
                     const float g = float(nb.vol * I.w_norm * I.dw_fac) * kshape_dw<float>(kt, q) / q;
                     const float wr = float(I.w_norm * inv_w_ave) * shp;
                     const float f = 1.0f + 0.2f * ((wr * wr) * (wr * wr));
                     sx = sx + f * (g * -xs);
                     sy = sy + f * (g * -ys);
                     float a[5];
                     basis5s(xs, ys, a);
                     S.add(vwn * shp, a, df);
                   }
                 });
  // shift vector, clamp (shift_interpolate_mod.f90:333-346)
  const double umax = __longlong_as_double((long long)A.umax[ii]);
  const double coe = I.pst_c * umax * I.dt * I.h;
  double dxs = -coe * double(sx), dys = -coe * double(sy);
  const double rs = sqrt(dxs * dxs + dys * dys);
  if (rs > 0.2 * I.h) {
    dxs = (0.2 * I.h) * dxs / rs;
    dys = (0.2 * I.h) * dys / rs;
  }
  // LS correction (shift_interpolate_mod.f90:352-415)
  double X[NF][5];
  if (cnt >= 5) {
    double M[5][5];
    unpack(S, M, X);
    lu_solve<5, NF>(M, X);
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
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) s += a[i] * X[f][i];
    acc[f] = s;
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
  A.rout[p] = out;
}

// Max particle speed per instance (PST_mod.f90:49-53).
__global__ void k_umax(int n, const Rec *__restrict__ rec, const int32_t *__restrict__ info,
                       unsigned long long *__restrict__ umax) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  int ii = 0;
  if (p < n) {
    const int32_t inf = info[p];
    ii = info_inst(inf);
    if (info_ptype(inf) == 0) {
      const Rec &r = rec[p];
      s = sqrt(r.u * r.u + r.v * r.v);
    }
  }
  const int i0 = __shfl(ii, 0);
  if (__all(ii == i0)) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s = fmax(s, __shfl_xor(s, o));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&umax[i0], s);
  } else if (p < n) {
    atomic_max_nonneg(&umax[ii], s);
  }
}

// ---------------------------------------------------------------------------
// Ghost walls (ghost_mod.f90:51-76)
// ---------------------------------------------------------------------------
struct GhostArgs {
  Rec *__restrict__ rec;  // current buffer: fluid read, walls written
  const int32_t *__restrict__ info;
  const int32_t *__restrict__ cstart;
  const int32_t *__restrict__ items;
  const InstDev *__restrict__ inst;
  const double *__restrict__ vm_x, *__restrict__ vm_y, *__restrict__ vm_nx, *__restrict__ vm_ny;
  const int32_t *__restrict__ vm_ptype, *__restrict__ vm_inst, *__restrict__ vm_cell, *__restrict__ vm_wall;
  double *vm_u, *vm_v, *vm_pre, *vm_tem;
  Options opt;
  int n_ext;
};

// Hydrostatic pressure of an isothermal wall side (extrapolate_VM_to_WL_mod
// .f90:57-61, 105-109, 209-213, 250-254).
__device__ __forceinline__ double wall_pre(const InstDev &I, bool has_t, int iso, double tw, double vpre,
                                           double wall_y, double vm_y) {
  if (has_t && iso == 1)
    return vpre + I.rho_ref * I.alpha_ref * (tw - I.tem_ave) * I.gravity * (wall_y - vm_y);
  return vpre;
}

template <int Q>
__device__ __forceinline__ void interp_lssph_b(const GhostArgs &A, const InstDev &I, int k, int kt, bool has_t,
                                               double (&val)[4]) {
  // interpolate_LSSPH_B_mod.f90:24-178: a = [1, x/h, y/h (, x^2/2h^2, xy/h^2, y^2/2h^2)]
  double M[Q * (Q + 1) / 2];
  double b[4][Q];
#pragma unroll
  for (int i = 0; i < Q * (Q + 1) / 2; ++i) M[i] = 0.0;
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int i = 0; i < Q; ++i) b[f][i] = 0.0;
  const double mx = A.vm_x[k], my = A.vm_y[k];
  const double inv_h = I.inv_h, inv_h2 = I.inv_h_sq;
  int num = 0;
  visit_cells(I, A.vm_cell[k] - I.cell_off, A.cstart, [&](int t0, int t1) {
    for (int t = t0; t < t1; ++t) {
      const int j = A.items[t];
      if (info_ptype(A.info[j]) != 0) continue;
      const Rec nb = A.rec[j];
      const double xj = nb.x - mx, yj = nb.y - my;
      const double r2 = sq_dist(xj, yj);
      if (r2 > I.h_eff_sq) continue;
      const double w = I.w_norm * kshape_w<double>(kt, sqrt(r2) * I.inv_h);
      ++num;
      const double vw = nb.vol * w;
      double a[Q];
      a[0] = 1.0;
      a[1] = xj * inv_h;
      a[2] = yj * inv_h;
      if constexpr (Q == 6) {
        a[3] = xj * xj * 0.5 * inv_h2;
        a[4] = xj * yj * inv_h2;
        a[5] = yj * yj * 0.5 * inv_h2;
      }
      const double fv[4] = {nb.u, nb.v, nb.pre, has_t ? nb.tem : 0.0};
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const double vwa = vw * a[i];
#pragma unroll
        for (int f = 0; f < 4; ++f) b[f][i] += vwa * fv[f];
#pragma unroll
        for (int r = 0; r <= i; ++r) M[i * (i + 1) / 2 + r] += vwa * a[r];
      }
    }
  });
  if (num >= Q) {
    double Am[Q][Q];
#pragma unroll
    for (int i = 0; i < Q; ++i)
#pragma unroll
      for (int r = 0; r <= i; ++r) {
        Am[r][i] = M[i * (i + 1) / 2 + r];
        Am[i][r] = Am[r][i];
      }
    lu_solve<Q, 4>(Am, b);
#pragma unroll
    for (int f = 0; f < 4; ++f) val[f] = b[f][0];
  } else {
#pragma unroll
    for (int f = 0; f < 4; ++f) val[f] = 0.0;
  }
}

__device__ __forceinline__ void interp_csph(const GhostArgs &A, const InstDev &I, int k, int kt, bool has_t,
                                            double (&val)[4]) {
  // interpolate_CSPH_mod.f90:24-130
  const double mx = A.vm_x[k], my = A.vm_y[k];
  double su = 0, sv = 0, sp = 0, st = 0, sw = 0;
  int num = 0;
  visit_cells(I, A.vm_cell[k] - I.cell_off, A.cstart, [&](int t0, int t1) {
    for (int t = t0; t < t1; ++t) {
      const int j = A.items[t];
      if (info_ptype(A.info[j]) != 0) continue;
      const Rec nb = A.rec[j];
      const double xij = mx - nb.x, yij = my - nb.y;
      const double r2 = sq_dist(xij, yij);
      if (r2 > I.h_eff_sq) continue;
      const double w = I.w_norm * kshape_w<double>(kt, sqrt(r2) * I.inv_h);
      ++num;
      const double vw = nb.vol * w;
      sw = sw + vw;
      su = su + vw * nb.u;
      sv = sv + vw * nb.v;
      sp = sp + vw * nb.pre;
      if (has_t) st = st + vw * nb.tem;
    }
  });
  if (num > 0 && sw > 1.0e-12) {
    val[0] = su / sw;
    val[1] = sv / sw;
    val[2] = sp / sw;
    val[3] = st / sw;
  } else {
    val[0] = val[1] = val[2] = val[3] = 0.0;
  }
}

__global__ void __launch_bounds__(256) k_ghost(GhostArgs A) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.n_ext) return;
  const Options &o = A.opt;
  const InstDev &I = A.inst[A.vm_inst[k]];
  const bool has_t = o.has_tem != 0;
  double val[4];
  if (o.wall_model == 1)
    interp_csph(A, I, k, o.kernel_type, has_t, val);
  else if (o.wall_model == 3)
    interp_lssph_b<6>(A, I, k, o.kernel_type, has_t, val);
  else
    interp_lssph_b<3>(A, I, k, o.kernel_type, has_t, val);
  A.vm_u[k] = val[0];
  A.vm_v[k] = val[1];
  A.vm_pre[k] = val[2];
  if (has_t) A.vm_tem[k] = val[3];

  Rec &w = A.rec[A.vm_wall[k]];
  const int t = A.vm_ptype[k];
  const double vu = val[0], vv = val[1], vp = val[2], vt = val[3];
  if (o.target_problem == 1) {
    // extrapolate_diff_test_mod.f90:38-104
    double u;
    const double vx = A.vm_x[k], vy = A.vm_y[k];
    if (t == 1)
      u = (o.u_bc_bottom == 1) ? vu : 2.0 * (sin(kPi * (vx - I.wl_thick))) - vu;
    else if (t == 2)
      u = (o.u_bc_top == 1) ? vu : 2.0 * (-sin(kPi * (vx - I.wl_thick))) - vu;
    else if (t == 3)
      u = (o.u_bc_left == 1) ? vu + kPi * cos(kPi * (vy - I.wl_thick)) * (w.x - vx) : -vu;
    else if (t == 4)
      u = (o.u_bc_right == 1) ? vu + (-kPi) * cos(kPi * (vy - I.wl_thick)) * (w.x - vx) : -vu;
    else
      u = -vu;
    w.u = u;
    return;
  }
  // extrapolate_VM_to_WL_mod.f90:40-273
  const double nx = A.vm_nx[k], ny = A.vm_ny[k];
  if (t >= 1 && t <= 4) {
    const int bc = (t == 1) ? o.u_bc_bottom : (t == 2) ? o.u_bc_top : (t == 3) ? o.u_bc_left : o.u_bc_right;
    if (bc == 1) {
      const double un = vu * nx + vv * ny;
      w.u = vu - 2.0 * un * nx;
      w.v = vv - 2.0 * un * ny;
    } else if (bc == 2) {
      w.u = -vu;
      w.v = -vv;
    } else if (bc == 3 && t == 2) {
      w.u = 2.0 * I.u_top - vu;
      w.v = -vv;
    }
  }
  const double vmy = A.vm_y[k];
  switch (t) {
    case 1:
      w.pre = wall_pre(I, has_t, o.tem_bc_bottom, I.tem_bottom, vp, w.y, vmy);
      if (has_t) w.tem = (o.tem_bc_bottom == 1) ? 2.0 * I.tem_bottom - vt : vt;
      break;
    case 2:
      w.pre = wall_pre(I, has_t, o.tem_bc_top, I.tem_top, vp, w.y, vmy);
      if (has_t) w.tem = (o.tem_bc_top == 1) ? 2.0 * I.tem_top - vt : vt;
      break;
    case 3:
      w.pre = vp;
      if (has_t) w.tem = (o.tem_bc_left == 1) ? 2.0 * I.tem_left - vt : vt;
      break;
    case 4:
      w.pre = vp;
      if (has_t) w.tem = (o.tem_bc_right == 1) ? 2.0 * I.tem_right - vt : vt;
      break;
    case 5:
    case 6:
      w.u = -vu;
      w.v = -vv;
      w.pre = wall_pre(I, has_t, o.tem_bc_bottom, I.tem_bottom, vp, w.y, vmy);
      if (has_t) w.tem = (o.tem_bc_bottom == 1) ? 2.0 * I.tem_bottom - vt : vt;
      break;
    default:  // 7, 8
      if (o.u_bc_top == 3) {
        w.u = 2.0 * I.u_top - vu;
        w.v = -vv;
      } else {
        w.u = -vu;
        w.v = -vv;
      }
      w.pre = wall_pre(I, has_t, o.tem_bc_top, I.tem_top, vp, w.y, vmy);
      if (has_t) w.tem = (o.tem_bc_top == 1) ? 2.0 * I.tem_top - vt : vt;
      break;
  }
}

// Virtual markers from the wall positions (init_virtual_markers_mod.f90:
// 61-137) and their fixed cells (init_cell_mod.f90:166-176).
__global__ void k_vm_init(int n_int, int n_ext, const Rec *__restrict__ rec, const int32_t *__restrict__ info,
                          const int32_t *__restrict__ phys_of, const InstDev *__restrict__ inst, double *vm_x,
                          double *vm_y, double *vm_nx, double *vm_ny, int32_t *vm_ptype, int32_t *vm_inst,
                          int32_t *vm_cell, int32_t *vm_wall, int32_t *escaped) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_ext) return;
  const int pw = phys_of[n_int + k];
  const int32_t inf = info[pw];
  const int t = info_ptype(inf), ii = info_inst(inf);
  const InstDev &I = inst[ii];
  const double wx = rec[pw].x, wy = rec[pw].y;
  const double r2 = 1.0 / sqrt(2.0);
  double x = wx, y = wy, nx = 0.0, ny = 0.0;
  switch (t) {
    case 1: y = 2.0 * I.y_min - wy; ny = -1.0; break;
    case 2: y = 2.0 * I.y_max - wy; ny = 1.0; break;
    case 3: x = 2.0 * I.x_min - wx; nx = -1.0; break;
    case 4: x = 2.0 * I.x_max - wx; nx = 1.0; break;
    case 5: x = 2.0 * I.x_min - wx; y = 2.0 * I.y_min - wy; nx = -r2; ny = -r2; break;
    case 6: x = 2.0 * I.x_max - wx; y = 2.0 * I.y_min - wy; nx = r2; ny = -r2; break;
    case 7: x = 2.0 * I.x_min - wx; y = 2.0 * I.y_max - wy; nx = -r2; ny = r2; break;
    default: x = 2.0 * I.x_max - wx; y = 2.0 * I.y_max - wy; nx = r2; ny = r2; break;
  }
  int cx = (int)ceil(x / I.h_eff) - 1, cy = (int)ceil(y / I.h_eff) - 1;
  if (t < 1 || t > 8 || cx < 0 || cx >= I.nx || cy < 0 || cy >= I.ny) {
    atomicAdd(escaped, 1);
    cx = cx < 0 ? 0 : (cx >= I.nx ? I.nx - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= I.ny ? I.ny - 1 : cy);
  }
  vm_x[k] = x;
  vm_y[k] = y;
  vm_nx[k] = nx;
  vm_ny[k] = ny;
  vm_ptype[k] = t;
  vm_inst[k] = ii;
  vm_cell[k] = I.cell_off + cx + cy * I.nx;
  vm_wall[k] = pw;
}

// ---------------------------------------------------------------------------
// Neighbor sets for validation (LSSPH_A_mod.f90:102-124)
// ---------------------------------------------------------------------------
__global__ void k_nbr_count(int n, const Rec *__restrict__ rec, const int32_t *__restrict__ info,
                            const int32_t *__restrict__ key, const int32_t *__restrict__ cstart,
                            const int32_t *__restrict__ items, const InstDev *__restrict__ inst,
                            const int32_t *__restrict__ orig_of, int32_t *counts_orig) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || info_ptype(info[p]) != 0) return;
  const InstDev &I = inst[info_inst(info[p])];
  const double2 me = load_xy(rec, p);
  int cnt = 0;
  visit_cells(I, key[p] - I.cell_off, cstart, [&](int t0, int t1) {
    for (int t = t0; t < t1; ++t) {
      const int j = items[t];
      if (j == p) continue;
      const double2 xy = load_xy(rec, j);
      const double dx = pdelta(xy.x, me.x, I.len_x, I.periodic), dy = pdelta(xy.y, me.y, I.len_y, I.periodic);
      if (sq_dist(dx, dy) <= I.h_eff_sq) ++cnt;
    }
  });
  counts_orig[orig_of[p]] = cnt;
}

__global__ void k_nbr_fill(int n, const Rec *__restrict__ rec, const int32_t *__restrict__ info,
                           const int32_t *__restrict__ key, const int32_t *__restrict__ cstart,
                           const int32_t *__restrict__ items, const InstDev *__restrict__ inst,
                           const int32_t *__restrict__ orig_of, const int64_t *__restrict__ offsets,
                           int32_t *list) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || info_ptype(info[p]) != 0) return;
  const InstDev &I = inst[info_inst(info[p])];
  const double2 me = load_xy(rec, p);
  int64_t w = offsets[orig_of[p]];
  visit_cells(I, key[p] - I.cell_off, cstart, [&](int t0, int t1) {
    for (int t = t0; t < t1; ++t) {
      const int j = items[t];
      if (j == p) continue;
      const double2 xy = load_xy(rec, j);
      const double dx = pdelta(xy.x, me.x, I.len_x, I.periodic), dy = pdelta(xy.y, me.y, I.len_y, I.periodic);
      if (sq_dist(dx, dy) <= I.h_eff_sq) list[w++] = orig_of[j];
    }
  });
}

// ---------------------------------------------------------------------------
// Steady-state check (check_steady_state_mod.f90:33-67)
// ---------------------------------------------------------------------------
__global__ void k_dfmax(int n, const Rec *__restrict__ rec, const int32_t *__restrict__ info,
                        const double *__restrict__ f_prev, unsigned long long *dfmax) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  int ii = 0;
  if (p < n) {
    ii = info_inst(info[p]);
    if (info_ptype(info[p]) == 0) s = fabs(f_prev[p] - rec[p].u);
  }
  const int i0 = __shfl(ii, 0);
  if (__all(ii == i0)) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s = fmax(s, __shfl_xor(s, o));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&dfmax[i0], s);
  } else if (p < n) {
    atomic_max_nonneg(&dfmax[ii], s);
  }
}

__global__ void k_steady_copy(int n, const Rec *__restrict__ rec, const int32_t *__restrict__ info,
                              const InstDev *__restrict__ inst, const unsigned long long *__restrict__ dfmax,
                              double *f_prev, int32_t *steady, int n_inst) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n_inst) {
    const double df = __longlong_as_double((long long)dfmax[p]);
    steady[p] = (df < inst[p].threshold) ? 1 : 0;
  }
  if (p >= n || info_ptype(info[p]) != 0) return;
  const int ii = info_inst(info[p]);
  const double df = __longlong_as_double((long long)dfmax[ii]);
  if (!(df < inst[ii].threshold)) f_prev[p] = rec[p].u;
}

// ---------------------------------------------------------------------------
// Upload / download between upload order and physical order
// ---------------------------------------------------------------------------
__global__ void k_download(int n, const Rec *__restrict__ rec, const int32_t *__restrict__ info,
                           const int32_t *__restrict__ phys_of, const int32_t *__restrict__ key,
                           const InstDev *__restrict__ inst, double *stage, int32_t *ptype_out,
                           int32_t *cell_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = phys_of[i];
  const Rec r = rec[p];
  stage[0 * (size_t)n + i] = r.x;
  stage[1 * (size_t)n + i] = r.y;
  stage[2 * (size_t)n + i] = r.u;
  stage[3 * (size_t)n + i] = r.v;
  stage[4 * (size_t)n + i] = r.rho;
  stage[5 * (size_t)n + i] = r.pre;
  stage[6 * (size_t)n + i] = r.tem;
  ptype_out[i] = info_ptype(info[p]);
  cell_out[i] = key[p] - inst[info_inst(info[p])].cell_off + 1;
}

struct RhsPtrs {
  const double *p[4];
};

__global__ void k_download_rhs(int n_int, const int32_t *__restrict__ phys_of, RhsPtrs rhs, double *stage) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_int) return;
  const int p = phys_of[i];
#pragma unroll
  for (int f = 0; f < 4; ++f) stage[f * (size_t)n_int + i] = rhs.p[f][p];
}

// Records in upload order (physical = upload order until the first sort).
__global__ void k_upload(int n, const double *__restrict__ stage, const int32_t *__restrict__ ptype,
                         const int32_t *__restrict__ inst_of, const InstDev *__restrict__ inst, Rec *rec,
                         int32_t *info, int has_tem) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Rec r;
  r.x = stage[0 * (size_t)n + i];
  r.y = stage[1 * (size_t)n + i];
  r.u = stage[2 * (size_t)n + i];
  r.v = stage[3 * (size_t)n + i];
  r.rho = stage[4 * (size_t)n + i];
  r.pre = stage[5 * (size_t)n + i];
  r.tem = has_tem ? stage[6 * (size_t)n + i] : 0.0;
  const int ii = inst_of[i];
  r.vol = inst[ii].sp_mass / r.rho;
  rec[i] = r;
  info[i] = make_info(ptype[i], ii);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline int blocks(int n, int b) { return (n + b - 1) / b; }

template <int CLS, bool MIXED>
static void lsa_launch(hipStream_t s, const LsaArgs &a, int block, size_t lds) {
  hipLaunchKernelGGL((k_lsa<CLS, MIXED>), dim3(blocks(a.n, block)), dim3(block), lds, s, a);
}

static int problem_class(const Options &o) {
  return o.target_problem == 1 ? 0 : (o.has_tem ? 2 : 1);
}

void launch_lsa(hipStream_t s, Device &d, int cur, StageMode mode) {
  LsaArgs a;
  a.rin = d.rec[cur];
  a.rout = d.rec[cur ^ 1];
  a.info = d.info;
  a.key = d.key;
  a.cstart = d.cell_start;
  a.items = d.items;
  a.inst = d.inst;
  for (int f = 0; f < 6; ++f) {
    a.old_[f] = d.old_[f];
    a.new_[f] = d.new_[f];
  }
  for (int f = 0; f < 4; ++f) a.rhs[f] = d.rhs[f];
  a.stats64 = d.stats64;
  a.counters = d.counters;
  a.n = d.n_total;
  a.kt = d.opt.kernel_type;
  a.list_len = d.list_len;
  a.mode = mode;
  const size_t lds = sizeof(int32_t) * (size_t)d.list_len * d.block;
  const bool mixed = d.opt.precision == LSSPH_MIXED;
  switch (problem_class(d.opt)) {
    case 0: mixed ? lsa_launch<0, true>(s, a, d.block, lds) : lsa_launch<0, false>(s, a, d.block, lds); break;
    case 1: mixed ? lsa_launch<1, true>(s, a, d.block, lds) : lsa_launch<1, false>(s, a, d.block, lds); break;
    default: mixed ? lsa_launch<2, true>(s, a, d.block, lds) : lsa_launch<2, false>(s, a, d.block, lds); break;
  }
}

void launch_umax(hipStream_t s, Device &d, int cur) {
  launch_fill_u64(s, d.umax, 0ull, d.n_inst);
  hipLaunchKernelGGL(k_umax, dim3(blocks(d.n_total, 256)), dim3(256), 0, s, d.n_total, d.rec[cur], d.info,
                     d.umax);
}

void launch_pst(hipStream_t s, Device &d, int cur) {
  PstArgs a;
  a.rin = d.rec[cur];
  a.rout = d.rec[cur ^ 1];
  a.info = d.info;
  a.key = d.key;
  a.cstart = d.cell_start;
  a.items = d.items;
  a.inst = d.inst;
  a.umax = d.umax;
  a.stats64 = d.stats64;
  a.n = d.n_total;
  a.kt = d.opt.kernel_type;
  a.list_len = d.list_len;
  a.has_tem = d.opt.has_tem;
  a.mixed = d.opt.precision;
  const size_t lds = sizeof(int32_t) * (size_t)d.list_len * d.block;
  const dim3 g(blocks(d.n_total, d.block)), b(d.block);
  const bool mixed = d.opt.precision == LSSPH_MIXED;
  if (d.opt.has_tem) {
    if (mixed)
      hipLaunchKernelGGL((k_pst<true, true>), g, b, lds, s, a);
    else
      hipLaunchKernelGGL((k_pst<true, false>), g, b, lds, s, a);
  } else {
    if (mixed)
      hipLaunchKernelGGL((k_pst<false, true>), g, b, lds, s, a);
    else
      hipLaunchKernelGGL((k_pst<false, false>), g, b, lds, s, a);
  }
}

void launch_ghost(hipStream_t s, Device &d, int cur) {
  if (d.n_ext == 0) return;
  GhostArgs a;
  a.rec = d.rec[cur];
  a.info = d.info;
  a.cstart = d.cell_start;
  a.items = d.items;
  a.inst = d.inst;
  a.vm_x = d.vm_x;
  a.vm_y = d.vm_y;
  a.vm_nx = d.vm_nx;
  a.vm_ny = d.vm_ny;
  a.vm_ptype = d.vm_ptype;
  a.vm_inst = d.vm_inst;
  a.vm_cell = d.vm_cell;
  a.vm_wall = d.vm_wall;
  a.vm_u = d.vm_u;
  a.vm_v = d.vm_v;
  a.vm_pre = d.vm_pre;
  a.vm_tem = d.vm_tem;
  a.opt = d.opt;
  a.n_ext = d.n_ext;
  hipLaunchKernelGGL(k_ghost, dim3(blocks(d.n_ext, 128)), dim3(128), 0, s, a);
}

void launch_vm_init(hipStream_t s, Device &d, int cur) {
  if (d.n_ext == 0) return;
  hipLaunchKernelGGL(k_vm_init, dim3(blocks(d.n_ext, 256)), dim3(256), 0, s, d.n_int, d.n_ext, d.rec[cur],
                     d.info, d.phys_of, d.inst, d.vm_x, d.vm_y, d.vm_nx, d.vm_ny, d.vm_ptype, d.vm_inst,
                     d.vm_cell, d.vm_wall, d.counters + kEscaped);
}

void launch_neighbor_count(hipStream_t s, Device &d, int cur, int32_t *counts_orig) {
  hipLaunchKernelGGL(k_nbr_count, dim3(blocks(d.n_total, 256)), dim3(256), 0, s, d.n_total, d.rec[cur], d.info,
                     d.key, d.cell_start, d.items, d.inst, d.orig_of, counts_orig);
}

void launch_neighbor_fill(hipStream_t s, Device &d, int cur, const int64_t *offsets_orig, int32_t *list) {
  hipLaunchKernelGGL(k_nbr_fill, dim3(blocks(d.n_total, 256)), dim3(256), 0, s, d.n_total, d.rec[cur], d.info,
                     d.key, d.cell_start, d.items, d.inst, d.orig_of, offsets_orig, list);
}

void launch_steady(hipStream_t s, Device &d, int cur) {
  launch_fill_u64(s, d.dfmax, 0ull, d.n_inst);
  hipLaunchKernelGGL(k_dfmax, dim3(blocks(d.n_total, 256)), dim3(256), 0, s, d.n_total, d.rec[cur], d.info,
                     d.f_prev, d.dfmax);
  const int n = d.n_total > d.n_inst ? d.n_total : d.n_inst;
  hipLaunchKernelGGL(k_steady_copy, dim3(blocks(n, 256)), dim3(256), 0, s, d.n_total, d.rec[cur], d.info,
                     d.inst, d.dfmax, d.f_prev, d.steady, d.n_inst);
}

void launch_download(hipStream_t s, Device &d, int cur) {
  const int n = d.n_total;
  hipLaunchKernelGGL(k_download, dim3(blocks(n, 256)), dim3(256), 0, s, n, d.rec[cur], d.info, d.phys_of, d.key,
                     d.inst, d.stage_f64, d.stage_i32, d.perm_tmp);
}

void launch_download_rhs(hipStream_t s, Device &d) {
  RhsPtrs r;
  for (int f = 0; f < 4; ++f) r.p[f] = d.rhs[f];
  hipLaunchKernelGGL(k_download_rhs, dim3(blocks(d.n_int, 256)), dim3(256), 0, s, d.n_int, d.phys_of, r,
                     d.stage_f64);
}

void launch_upload(hipStream_t s, Device &d, const double *stage, const int32_t *ptype, const int32_t *inst_of,
                   int has_tem) {
  const int n = d.n_total;
  hipLaunchKernelGGL(k_upload, dim3(blocks(n, 256)), dim3(256), 0, s, n, stage, ptype, inst_of, d.inst,
                     d.rec[0], d.info, has_tem);
}

}  // namespace lssph
