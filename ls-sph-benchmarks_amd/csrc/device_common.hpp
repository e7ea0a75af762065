// device_common.hpp -- data layout and device helpers shared by the LS-SPH
// kernels for gfx950.
//
// Particle records live in HBM as 64-byte AoS records (one per particle) in
// cell order ("slots"), double-buffered across passes: a pair kernel reads
// every neighbor from buffer A and writes the updated particle to the same
// slot of buffer B, so the stage update is fused into the operator's
// epilogue and the cell order survives until the next cell build.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lssph {

constexpr double kPi = 3.141592653589793;  // acos(-1.0d0)

// One particle; 64 B = four 16-B chunks.  vol = sp_mass / rho is kept so
// pair loops never divide (the quotient is the same IEEE value the reference
// forms per pair, LSSPH_A_mod.f90:136).  The order puts what one neighbor
// gather needs in as few chunks as possible: (x, y) and (u, v); rho and vol
// side by side for the shifting (rho differences, volumes), vol and pre side
// by side for the LS operator (one 16-B load at byte 40).
struct alignas(64) Rec {
  double x, y, u, v;
  double rho, vol, pre, tem;
};

// The neighbor fields of the mixed-precision path: (u, v, pre, vol) of a
// record rounded to FP32, 16 B per slot in an array beside the records (the
// same slot order and double buffering).  A mixed pair walk gathers these
// instead of two 16-B chunks of the 64-B record: one load per neighbor, and
// eight neighbors per 128-B cache line instead of two.
__device__ __forceinline__ float4 nb_fields(const Rec &r) {
  return make_float4(float(r.u), float(r.v), float(r.pre), float(r.vol));
}

// Writes record r to slot s, and its neighbor fields when the handle keeps
// them (nbf non-null: the mixed path).
__device__ __forceinline__ void store_rec(Rec *__restrict__ out, float4 *__restrict__ nbf, int s, const Rec &r) {
  out[s] = r;
  if (nbf) nbf[s] = nb_fields(r);
}

// Packed per-particle info: ptype in bits 0..7 (0 = fluid), instance id in
// bits 8..31.
__host__ __device__ inline int info_ptype(int32_t i) { return i & 0xff; }
__host__ __device__ inline int info_inst(int32_t i) { return i >> 8; }
__host__ __device__ inline int32_t make_info(int ptype, int inst) { return ptype | (inst << 8); }

// Per-instance constants derived once from lssph_params_t.
struct InstDev {
  double h, inv_h, inv_h2, h_eff_sq, inv_h_eff, h_eff;
  double inv_h_sq;    // inv_h * inv_h (type-B basis, interpolate_LSSPH_B_mod.f90:152)
  double inv_h3;      // inv_h2 / h (3rd-order Taylor terms, LSSPH_A_mod.f90:148-153)
  double inv_h4;      // inv_h3 / h (4th-order terms, :155-162)
  double threshold;   // steady-state threshold
  double sp_mass, w_ave, wl_thick, len_x, len_y;
  double inv_cw_x, inv_cw_y;  // internal cell grid: 1 / cell width (>= h_eff + skin)
  double inv_rcw_x, inv_rcw_y;  // periodic reference grid: cells / length
  double skin_half_sq;        // (0.45 skin)^2: largest squared move between cell builds
  double w_norm;      // kernel normalisation, e.g. 7/(4 pi h^2) for Wendland C2
  double dw_fac;      // dW/dr = w_norm * shape' * dw_fac
  double rho_ref, vis_ref, c_sq, thermal_dif, alpha_ref, gravity, tem_ave;
  double zeta_sq, xi_sq, delta;  // delta = coe_delta_sph * c * h * zeta
  double pst_c, dt, u_top;
  double dt_half, dt_third, dt_sixth;  // dt / 2, dt / 3, dt / 6 (RK stage weights, formed once)
  double tem_top, tem_bottom, tem_left, tem_right;
  double diff_coef, diff_source, cond_max;
  double x_min, x_max, y_min, y_max;  // fluid box (walls beyond)
  int32_t nx, ny, cell_off, periodic;  // internal cell grid
  int32_t rnx, rny;                     // the reference's cell grid (cell_idx output)
  int32_t retired;  // isolation mode: set (by a vector store) when the instance is retired
  // Cell columns per stencil reach in x: the internal cells are (h_eff +
  // skin) / kx wide in x and h_eff + skin tall, and the stencil of a cell is
  // the 2 kx + 1 columns of the rows above, at and below it.
  int32_t kx;
};

// Model options shared by every instance of a handle.
struct Options {
  int32_t target_problem, kernel_type, wall_model, rk;
  int32_t u_bc_top, u_bc_bottom, u_bc_left, u_bc_right;
  int32_t tem_bc_top, tem_bc_bottom, tem_bc_left, tem_bc_right;
  int32_t has_tem, precision;
  int32_t delta;      // any instance with coe_delta_sph != 0
  int32_t sph_model;  // 1-2 classical SPH, 3-6 LS type A with q = 5, 9, 14, 20
};

// Hardware square root and reciprocal of FP32 (1 ulp), for the mixed path's
// FP32 pair terms, whose accuracy is set by the FP32 sums anyway.
__device__ __forceinline__ float fast_sqrtf(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// |(xs, ys)| of the mixed path's scaled offsets with the fused multiply-add
// written out: left to FMA contraction, xs * xs + ys * ys may fuse either
// product, and the choice can differ between the inlined copies of a pair
// body (the walk has one per list layout), which would make a particle's
// FP32 sums depend on the path its wavefront took.
__device__ __forceinline__ float radius_s(float xs, float ys) {
  return __builtin_amdgcn_sqrtf(__builtin_fmaf(xs, xs, ys * ys));
}

// ---------------------------------------------------------------------------
// Kernel shapes (kernel_functions_mod.f90:26-146) as functions of q = r/h
// without the normalisation factor.
// ---------------------------------------------------------------------------
// KT > 0 fixes the kernel at compile time (the runtime kt is then ignored);
// KT = 0 selects it at run time.  INSIDE: the caller guarantees q <= 2 (a
// neighbor accepted by r^2 <= h_eff^2), so the Wendland C2 support test is
// left out (in FP32, q may round a few ulp above 2, where t^4 ~ 1e-29).
template <typename T, int KT = 0, bool INSIDE = false>
__device__ __forceinline__ T kshape_w(int kt_rt, T q) {
  const int kt = KT ? KT : kt_rt;
  const T one = T(1), half = T(0.5), two = T(2);
  if (kt == 3) {
    if (!INSIDE && q > two) return T(0);
    T t = one - half * q, t2 = t * t;
    return (t2 * t2) * (one + two * q);
  }
  if (kt == 1) {
    if (q <= one) {
      T a = two - q, b = one - q;
      return a * a * a - T(4) * (b * b * b);
    }
    if (q <= two) { T a = two - q; return a * a * a; }
    return T(0);
  }
  if (kt == 2) {
    T a = T(3) - q, b = two - q, c = one - q;
    T a5 = a * a * a * a * a;
    if (q <= one) return a5 - T(6) * (b * b * b * b * b) + T(15) * (c * c * c * c * c);
    if (q <= two) return a5 - T(6) * (b * b * b * b * b);
    if (q <= T(3)) return a5;
    return T(0);
  }
  if (kt == 4) {
    if (q > two) return T(0);
    T t = one - half * q, t2 = t * t;
    return (t2 * t2 * t2) * (one + T(3) * q + T(35) * (q * q) / T(12));
  }
  // kt == 5
  if (q > two) return T(0);
  T t = one - half * q, t2 = t * t, t4 = t2 * t2;
  return (t4 * t4) * (one + T(4) * q + T(25) * (q * q) / T(4) + T(4) * (q * q * q));
}

template <typename T, int KT = 0>
__device__ __forceinline__ T kshape_dw(int kt_rt, T q) {
  const int kt = KT ? KT : kt_rt;
  const T one = T(1), half = T(0.5), two = T(2);
  if (kt == 3) {
    if (q > two) return T(0);
    T t = one - half * q, t3 = t * t * t;
    return -two * t3 * (one + two * q) + two * (t3 * t);
  }
  if (kt == 1) {
    if (q <= one) { T a = two - q, b = one - q; return a * a - T(4) * (b * b); }
    if (q <= two) { T a = two - q; return a * a; }
    return T(0);
  }
  if (kt == 2) {
    T a = T(3) - q, b = two - q, c = one - q;
    T a4 = a * a * a * a;
    if (q <= one) return a4 - T(6) * (b * b * b * b) + T(15) * (c * c * c * c);
    if (q <= two) return a4 - T(6) * (b * b * b * b);
    if (q <= T(3)) return a4;
    return T(0);
  }
  if (kt == 4) {
    if (q > two) return T(0);
    T t = one - half * q, t5 = t * t * t * t * t;
    return -T(3) * t5 * (one + T(3) * q + T(35) * (q * q) / T(12)) +
           one * (t5 * t) * (T(3) + T(35) * q / T(6));
  }
  if (q > two) return T(0);
  T t = one - half * q, t7 = t * t * t * t * t * t * t;
  return -T(4) * t7 * (one + T(4) * q + T(25) * (q * q) / T(4) + T(4) * (q * q * q)) +
         one * (t7 * t) * (T(4) + T(25) * q / T(2) + T(12) * (q * q));
}

// ---------------------------------------------------------------------------
// In-register LU with partial pivoting (Golub & Van Loan, Alg. 3.4.1) and
// DGESV's interface semantics: pivot = first row holding the largest |a| of
// the column; a zero pivot sets info and leaves the right-hand sides as they
// were.  All indices are compile-time after unrolling; row interchanges are
// selects, so nothing spills to scratch.
// ---------------------------------------------------------------------------
template <int N, int NR, typename T>
__device__ __forceinline__ int lu_factor(T (&a)[N][N], int (&piv)[N]) {
  int info = 0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    int p = j;
    T amax = fabs(a[j][j]);
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      T t = fabs(a[i][j]);
      bool gt = t > amax;
      amax = gt ? t : amax;
      p = gt ? i : p;
    }
    piv[j] = p;
    // swap rows j and p (whole rows, as dgetf2 does)
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      bool s = (p == i);
#pragma unroll
      for (int c = 0; c < N; ++c) {
        T aj = a[j][c], ai = a[i][c];
        a[j][c] = s ? ai : aj;
        a[i][c] = s ? aj : ai;
      }
    }
    T pv = a[j][j];
    if (pv == T(0)) {
      info = info ? info : j + 1;
    } else {
#pragma unroll
      for (int i = j + 1; i < N; ++i) a[i][j] = a[i][j] / pv;
    }
#pragma unroll
    for (int i = j + 1; i < N; ++i)
#pragma unroll
      for (int c = j + 1; c < N; ++c) a[i][c] -= a[i][j] * a[j][c];
  }
  return info;
}

// Solves with the factors of lu_factor: row interchanges, L (unit), U.
template <int N, int NR, typename T>
__device__ __forceinline__ void lu_apply(const T (&a)[N][N], const int (&piv)[N], T (&b)[NR][N]) {
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      bool s = (piv[j] == i);
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        T bj = b[r][j], bi = b[r][i];
        b[r][j] = s ? bi : bj;
        b[r][i] = s ? bj : bi;
      }
    }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int i = j + 1; i < N; ++i) b[r][i] -= b[r][j] * a[i][j];
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
      b[r][j] = b[r][j] / a[j][j];
#pragma unroll
      for (int i = 0; i < j; ++i) b[r][i] -= b[r][j] * a[i][j];
    }
  }
}

// DGESV: factor, then solve only when no pivot is exactly zero (otherwise
// the right-hand sides stay untouched and info = j > 0 is returned).
template <int N, int NR, typename T>
__device__ __forceinline__ int lu_solve(T (&a)[N][N], T (&b)[NR][N]) {
  int piv[N];
  int info = lu_factor<N, NR>(a, piv);
  if (info == 0) lu_apply<N, NR>(a, piv, b);
  return info;
}

// Instance constants through the constant address space: loads at a
// wave-uniform address (instance index from readfirstlane) become scalar
// loads instead of per-lane vector loads.  The only field written on the
// device is `retired` (isolation mode), by vector stores that later kernels
// see; a kernel running while it is set may read either value.
__device__ __forceinline__ const InstDev *inst_const(const InstDev *p, int i) {
  using CInst = const __attribute__((address_space(4))) InstDev;
  return (const InstDev *)((CInst *)p + i);
}

// XCD-aware block index: the hardware deals workgroups round-robin to the 8
// XCDs (block b runs on XCD b % 8), each with its own L2.  Remapping gives
// XCD k the contiguous logical blocks [k * G/8, (k+1) * G/8), so the slot
// ranges whose neighbor data overlap share one L2.  The grid must be a
// multiple of 8 blocks; logical blocks past the work return early.
constexpr int kNumXcd = 8;

__device__ __forceinline__ int xcd_block() {
  const int per = gridDim.x / kNumXcd;
  return (blockIdx.x % kNumXcd) * per + blockIdx.x / kNumXcd;
}

__host__ __device__ inline int xcd_grid(int n, int block) {
  const int g = (n + block - 1) / block;
  return (g + kNumXcd - 1) / kNumXcd * kNumXcd;
}

// Atomic max of a non-negative double through its bit pattern (IEEE order
// equals integer order for non-negative values).
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *addr, double v) {
  atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

}  // namespace lssph
