// operators.hip -- LS-SPH pair operators for gfx950.
//
// One thread per slot.  A pair kernel scans the stencil's candidates (x/y
// only, 16 B each) and compacts the accepted slots into a per-thread list in
// LDS (per-wave rows, bank-conflict free), then walks that list with every
// lane busy, loading the 48 B of each neighbor record it uses; the ~100-flop
// pair body thus stays out of the 1-in-3 divergent candidate test.  Full
// wavefronts inside one cell row stage their candidates' positions in LDS
// with coalesced loads and scan from there.  A thread whose list overflows
// walks the candidates again directly.
//
// The LS type-A kernel (LSSPH_A_mod.f90:26-271) assembles the q x q moment
// matrix and right-hand sides in registers, solves, forms the RHS of the
// governing equations and applies the RK2/RK4 stage update (RK2_mod.f90:
// 83-110, RK4_mod.f90:234-288) in its epilogue, writing the particle's new
// record to the same slot of the other record buffer.  The classical-SPH
// kernel (classical_SPH_mod.f90) shares the scan, the walk and the epilogue.
//   FP64 path: LU with partial pivoting (DGESV semantics).
//   Mixed path: FP32 pair sums on FP64-formed offsets and differences, FP64
//   LDL^T + one refinement step; particles whose LDL^T pivots are not
//   positive or whose moment matrix has a 1-norm condition number above
//   cond_max are queued and redone by the FP64 path in a second,
//   list-driven launch (k_lsa_fallback, k_pst_fallback).
#include "ls_common.hpp"

namespace lssph {

// Q = 5 (SPH_MODEL 3), 9 or 14 (SPH_MODEL 4-5, FP64; at q = 14 part of the
// sums live in scratch, and it is still 4x faster than a group of lanes per
// particle).  q = 20 runs on the lane-group kernels of highorder.hip.
template <int CLS, bool MIXED, bool DELTA, int KT, int Q>
__device__ __forceinline__ void lsa_particle(const LsaArgs &A, const InstDev &I, int s, int tid, int32_t *lds) {
  constexpr int NF = Layout<CLS, DELTA>::NF;
  const Rec me = A.mir[s];
  const int lc = local_cell(I, me.x, me.y);
  // list entries, the particle itself included
  NbList nl = make_list(lds, A.list_len, tid);
  build_list(A.mir, A.cstart, I, lc, me.x, me.y, I.h_eff_sq, nl, A.staged != 0);
  const int cnt = nl.cnt;
  if (cnt > A.list_len) {
    atomicAdd(&A.stats64[kOverflows], 1ull);
    atomicMax(&A.counters[kMaxNbr], cnt - 1);
  }
  double d[NF][5];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int i = 0; i < 5; ++i) d[f][i] = 0.0;
  if (cnt - 1 >= Q) {  // LSSPH_A_mod.f90:208 (num_neighbors >= q)
    if constexpr (MIXED) {
      if (!derivs32<CLS, DELTA, KT>(A, I, lc, s, me, nl, d)) {
        const int k = atomicAdd(A.fb_count, 1);
        A.fb_list[k] = s;
        return;
      }
    } else {
      derivs64<CLS, DELTA, KT, Q>(A, I, lc, s, me, nl, d);
    }
  }
  // Re-read the particle's own record for the epilogue instead of keeping it
  // live across the solve (the compiler barrier forbids reusing the first
  // load), which lowers the kernel's peak register count.
  __asm__ volatile("" ::: "memory");
  const Rec me2 = A.mir[s];
  finish<CLS, DELTA>(A, I, A.items[s], s, me2, d);
}

// The FP64 flow kernel without delta-SPH is held to 128 VGPRs, 4 waves per
// SIMD instead of 3 (2 registers spill; 6 % faster at 2048^2); the others
// would spill far more.
template <int CLS, bool MIXED, bool DELTA, int KT, int Q = 5>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(!MIXED && CLS == 1 && !DELTA && Q == 5 ? 4 : 1))) k_lsa(LsaArgs A) {
  static_assert(!MIXED || Q == 5, "the mixed path is q = 5");
  extern __shared__ int32_t lds[];
  const int tid = threadIdx.x, B = blockDim.x;
  const int s = xcd_block() * B + tid;
  if (s >= A.n) return;
  const int32_t inf = A.sinfo[s];
  // walls (and the particles of retired instances) move on with the records
  auto keep = [&] {
    if (A.mode != kRhsOnly) store_rec(A.mout, A.nbf_out, s, A.mir[s]);
    if (A.spd) A.spd[s] = 0.0;
  };
  if (info_ptype(inf) != 0) {
    keep();
    return;
  }
  for_instance(A.inst, inf, [&](const InstDev &I, int) { lsa_particle<CLS, MIXED, DELTA, KT, Q>(A, I, s, tid, lds); },
               keep);
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
    derivs64<CLS, DELTA, KT>(A, I, lc, s, me, no_list(), d);
    atomicAdd(&A.stats64[kFallbacks], 1ull);
    finish<CLS, DELTA>(A, I, A.items[s], s, me, d);
  }
  // The last block to finish empties the queue for the next launch: every
  // block has read the count before it reports, so no clearing kernel has
  // to run before the mixed kernel.
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(A.fb_done, 1) == (int)gridDim.x - 1) {
      *A.fb_count = 0;
      *A.fb_done = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// Classical SPH (classical_SPH_mod.f90:26-170), SPH_MODEL 1 (SUM: pressure
// gradient over p_j + p_i) and 2 (p_j - p_i), FP64, with the stage update of
// k_lsa.  The reference's x_ij = x_i - x_j is the exact negation of the
// offsets the walk passes; the kernel-gradient sums are kept as the
// reference forms them and handed to the epilogue as derivatives:
// d[U][0] = du/dx, d[V][1] = dv/dy, d[P][0..1] = grad p, d[*][2] = Laplacian.
// ---------------------------------------------------------------------------
template <int CLS, bool DELTA, bool SUM>
__device__ __forceinline__ void csph_particle(const LsaArgs &A, const InstDev &I, int s, int tid, int32_t *lds) {
  using L = Layout<CLS, DELTA>;
  constexpr int NF = L::NF;
  const Rec me = A.mir[s];
  const int lc = local_cell(I, me.x, me.y);
  NbList nl = make_list(lds, A.list_len, tid);
  build_list(A.mir, A.cstart, I, lc, me.x, me.y, I.h_eff_sq, nl, A.staged != 0);
  const int cnt = nl.cnt;
  if (cnt > A.list_len) {
    atomicAdd(&A.stats64[kOverflows], 1ull);
    atomicMax(&A.counters[kMaxNbr], cnt - 1);
  }
  const int kt = A.kt;
  double du_dx = 0.0, dv_dy = 0.0, dp_dx = 0.0, dp_dy = 0.0;
  double lap_u = 0.0, lap_v = 0.0, lap_t = 0.0, lap_rho = 0.0;
  walk_neighbors(A.mir, A.cstart, I, lc, s, me, nl,
                 [&](const Rec &nb, double dx, double dy, int) {
                   const double r = sqrt(sq_dist(dx, dy));
                   if (!(r > 0.0)) return;  // the particle itself
                   const double x_ij = -dx, y_ij = -dy;
                   const double dwdr = I.w_norm * kshape_dw<double, 0>(kt, r * I.inv_h) * I.dw_fac;
                   const double vj = nb.vol;
                   const double dwx = dwdr * x_ij / r * vj;  // :103-104
                   const double dwy = dwdr * y_ij / r * vj;
                   const double g = 2.0 * vj * dwdr / r;     // :124-134
                   du_dx = du_dx + dwx * (nb.u - me.u);
                   lap_u = lap_u + g * (me.u - nb.u);
                   if constexpr (CLS >= 1) {
                     dv_dy = dv_dy + dwy * (nb.v - me.v);
                     const double pp = SUM ? nb.pre + me.pre : nb.pre - me.pre;
                     dp_dx = dp_dx + dwx * pp;
                     dp_dy = dp_dy + dwy * pp;
                     lap_v = lap_v + g * (me.v - nb.v);
                   }
                   if constexpr (CLS == 2) lap_t = lap_t + g * (me.tem - nb.tem);
                   if constexpr (DELTA) lap_rho = lap_rho + g * (me.rho - nb.rho);
                 });
  double d[NF][5];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int i = 0; i < 5; ++i) d[f][i] = 0.0;
  d[L::U][0] = du_dx;
  d[L::U][2] = lap_u;
  if constexpr (CLS >= 1) {
    d[L::V][1] = dv_dy;
    d[L::V][2] = lap_v;
    d[L::P][0] = dp_dx;
    d[L::P][1] = dp_dy;
  }
  if constexpr (CLS == 2) d[L::T][2] = lap_t;
  if constexpr (DELTA) d[L::RHO][2] = lap_rho;
  __asm__ volatile("" ::: "memory");
  const Rec me2 = A.mir[s];
  finish<CLS, DELTA>(A, I, A.items[s], s, me2, d);
}


template <int CLS, bool DELTA, bool SUM>
__global__ void __launch_bounds__(256) k_csph(LsaArgs A) {
  extern __shared__ int32_t lds[];
  const int tid = threadIdx.x, B = blockDim.x;
  const int s = xcd_block() * B + tid;
  if (s >= A.n) return;
  const int32_t inf = A.sinfo[s];
  auto keep = [&] {
    if (A.mode != kRhsOnly) store_rec(A.mout, A.nbf_out, s, A.mir[s]);
    if (A.spd) A.spd[s] = 0.0;
  };
  if (info_ptype(inf) != 0) {
    keep();
    return;
  }
  for_instance(A.inst, inf, [&](const InstDev &I, int) { csph_particle<CLS, DELTA, SUM>(A, I, s, tid, lds); }, keep);
}

// ---------------------------------------------------------------------------
// Particle shifting (PST_mod.f90:37-114, shift_interpolate_mod.f90:23-300)
// ---------------------------------------------------------------------------

// Q = 5 for LS-SPH (SPH_MODEL 3), 9 and 14 for SPH_MODEL 4-5 (FP64), 2 for
// the classical models.  tid >= 0: the neighbors come from the LDS list the
// scan fills; tid < 0 (the FP64 redo of k_pst_fallback): from a direct walk
// over the stencil, in the same order.
template <bool T, bool MIXED, int KT, int Q>
__device__ __forceinline__ void pst_particle(const PstArgs &A, const InstDev &I, int ii, int s, int tid,
                                             int32_t *lds) {
  constexpr int NF = T ? 4 : 3;  // u, v, rho (, tem)
  using R = typename std::conditional<MIXED, float, double>::type;
  const Rec me = A.mir[s];
  const int lc = local_cell(I, me.x, me.y);
  const int kt = A.kt;
  // list entries, the particle itself included
  NbList nl = no_list();
  if (tid >= 0) {
    nl = make_list(lds, A.list_len, tid);
    build_list(A.mir, A.cstart, I, lc, me.x, me.y, I.h_eff_sq, nl, A.staged != 0);
    if (nl.cnt > A.list_len) atomicAdd(&A.stats64[kOverflows], 1ull);
  }
  // a particle redone in FP64 was queued with cnt - 1 >= Q
  const int cnt = nl.cnt;

  LsSums<R, NF, Q> S;
  S.zero();
  R sx = R(0), sy = R(0);
  const double inv_w_ave = 1.0 / I.w_ave;
  // the deep walk for the mixed Wendland kernel without temperature (the
  // others would spill under the 128-register cap)
  walk_neighbors<MIXED && !T && KT == 3>(A.mir, A.cstart, I, lc, s, me, nl,
                 [&](const Rec &nb, double dx, double dy, int) {
                   // the shifting gathers (u, v) and (rho, vol) from the
                   // record, one cache line per neighbor: its fit needs rho,
                   // which the FP32 neighbor fields do not hold
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
                     double a[Q];
                     basis_q<Q>(dx, dy, I, a);
                     S.add(vj * w, a, df);
                   } else {
                     const float xs = float(dx * I.inv_h), ys = float(dy * I.inv_h);
                     const float q = radius_s(xs, ys);
                     const float vol = float(nb.vol);
                     float shp, g;
                     if constexpr (KT == 3) {
                       // Wendland C2: shape = t^4 (1 + 2q), shape' = -5 q t^3 with
                       // t = 1 - q/2, so dW/dr * (-x/r) * V = V w_norm dw_fac
                       // (-5 t^3) (-xs): no division by q, and the particle
                       // itself (xs = ys = 0) adds nothing
                       const float t = 1.0f - 0.5f * q, t3 = t * t * t;
                       shp = t3 * t * (1.0f + 2.0f * q);  // q <= 2: an accepted neighbor
                       g = vol * float(I.w_norm * I.dw_fac) * (-5.0f * t3);
                     } else {
                       shp = kshape_w<float, KT>(kt, q);
                       // dW/dr * (-x/r) * V = V * w_norm * dw_fac * shape'(q) * (-xs / q)
                       const float g0 = vol * float(I.w_norm * I.dw_fac) * kshape_dw<float, KT>(kt, q) * fast_rcpf(q);
                       g = q > 0.0f ? g0 : 0.0f;  // the particle itself adds no shift
                     }
                     const float wr = float(I.w_norm * inv_w_ave) * shp;
                     const float f = 1.0f + 0.2f * ((wr * wr) * (wr * wr));
                     sx = sx + f * (g * -xs);
                     sy = sy + f * (g * -ys);
                     float a[5];
                     basis5s(xs, ys, a);
                     // the common factor w_norm of M and b cancels in the solve
                     S.add_reduced(vol * shp, a, df);
                   }
                 });
  if constexpr (MIXED) S.complete();
  {
  // The epilogue is evaluated without FMA contraction, like the
  // reference's (baseline x86-64 has no FMA), so that every kernel that
  // inlines it -- k_pst and its FP64 redo -- rounds it identically.
#pragma clang fp contract(off)
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
  double X[NF][Q];
  if (cnt - 1 >= Q) {  // shift_interpolate_mod.f90:240 (num_neighbors >= q)
    if constexpr (MIXED) {
      if (!ldlt_solve_refine<NF>(S, X, I.cond_max)) {
        // the whole particle -- shift sums and fit -- is redone in FP64 by
        // k_pst_fallback
        const int k = atomicAdd(A.fb_count, 1);
        A.fb_list[k] = s;
        return;
      }
    } else {
      double M[Q][Q];
      unpack(S, M, X);
      lu_solve<Q, NF>(M, X);
    }
  } else {
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int i = 0; i < Q; ++i) X[f][i] = 0.0;
  }
  double a[Q];
  basis_q<Q>(dxs, dys, I, a);
  double acc[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < Q; ++i) t += a[i] * X[f][i];
    acc[f] = t;
  }
  // the particle's own record is read again rather than kept live across
  // the solve (as in k_lsa), which lowers the peak register count
  __asm__ volatile("" ::: "memory");
  const Rec me2 = A.mir[s];
  Rec out = me2;
  const bool per = I.periodic;
  out.x = wrap_pos(me2.x + dxs, I.len_x, per);
  out.y = wrap_pos(me2.y + dys, I.len_y, per);
  out.u = me2.u + acc[0];
  out.v = me2.v + acc[1];
  out.rho = me2.rho + acc[2];
  if constexpr (T) out.tem = me2.tem + acc[3];
  out.pre = I.c_sq * (out.rho - I.rho_ref);
  out.vol = I.sp_mass / out.rho;
  store_rec(A.mout, A.nbf_out, s, out);
  record_move(A.mv, I, s, out.x, out.y);
  }
}


// The mixed kernel without temperature is held to 128 VGPRs (4 waves per
// SIMD); its two instance paths otherwise need 129.  With temperature the
// cap would spill ~50 registers, so that one keeps 3 waves.
template <bool T, bool MIXED, int KT, int Q>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MIXED && !T ? 4 : 1))) k_pst(PstArgs A) {
  static_assert(!MIXED || Q == 5, "the mixed path is q = 5");
  extern __shared__ int32_t lds[];
  const int tid = threadIdx.x, B = blockDim.x;
  const int s = xcd_block() * B + tid;
  if (s >= A.n) return;
  const int32_t inf = A.sinfo[s];
  auto keep = [&] { store_rec(A.mout, A.nbf_out, s, A.mir[s]); };  // walls move on with the records
  if (info_ptype(inf) != 0) {
    keep();
    return;
  }
  for_instance(A.inst, inf, [&](const InstDev &I, int ii) { pst_particle<T, MIXED, KT, Q>(A, I, ii, s, tid, lds); },
               keep);
}

// FP64 redo of the particles the mixed shifting kernel queued: shift sums,
// moment sums and solve in FP64 over a direct stencil walk (the order of the
// FP64 kernel's list), so a redone particle equals the FP64 path's result.
template <bool T, int KT>
__global__ void __launch_bounds__(256) k_pst_fallback(PstArgs A) {
  const int count = *A.fb_count;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const int s = A.fb_list[i];
    const int ii = info_inst(A.sinfo[s]);
    pst_particle<T, false, KT, 5>(A, A.inst[ii], ii, s, -1, nullptr);
    atomicAdd(&A.stats64[kFallbacks], 1ull);
  }
  // the last block to finish empties the queue (as in k_lsa_fallback)
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(A.fb_done, 1) == (int)gridDim.x - 1) {
      *A.fb_count = 0;
      *A.fb_done = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// Per-instance max reductions over the fluid particles of an instance (its
// particles occupy the contiguous slot range of its cells).
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

// The same maximum over the speeds the last RK stage stored (0 for walls):
// 8 B per slot instead of the records' u, v.
__global__ void __launch_bounds__(256) k_umax_spd(const double *__restrict__ spd, const int32_t *__restrict__ cstart,
                                                  const InstDev *__restrict__ inst, unsigned long long *umax) {
  const InstDev &I = inst[blockIdx.y];
  const int t0 = cstart[I.cell_off], t1 = cstart[I.cell_off + I.nx * I.ny];
  const int stride = gridDim.x * blockDim.x;
  double s = 0.0;
  int t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
  // four independent loads per iteration (a maximum is exact in any order)
  for (; t + 3 * stride < t1; t += 4 * stride)
    s = fmax(fmax(s, spd[t]), fmax(fmax(spd[t + stride], spd[t + 2 * stride]), spd[t + 3 * stride]));
  for (; t < t1; t += stride) s = fmax(s, spd[t]);
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
// Validation: the neighbor lists the pair kernels walk, as they walk them.
// The same scan (build_list) and walk (walk_neighbors, four-deep) as k_lsa
// and k_pst, with the same launch shape and LDS layout; the walk records
// the upload index of each neighbor it visits (-1 - index for a visit at a
// distance above h_eff, which never happens in a correct walk) and the path
// the particle took: 0 per-lane scan, 1 staged scan row by row, 2 staged
// scan with the positions kept in LDS, 3 list overflow (direct walk).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void pair_list_particle(const LsaArgs &A, const InstDev &I, int s, int tid, int32_t *lds,
                                                   int32_t *counts, int32_t *list, int32_t *path, int cap) {
  const Rec me = A.mir[s];
  const int lc = local_cell(I, me.x, me.y);
  NbList nl = make_list(lds, A.list_len, tid);
  build_list(A.mir, A.cstart, I, lc, me.x, me.y, I.h_eff_sq, nl, A.staged != 0);
  const int p = A.orig_of[A.items[s]];
  int k = 0;
  walk_neighbors<true>(A.mir, A.cstart, I, lc, s, me, nl, [&](const Rec &, double dx, double dy, int t) {
    if (t == s) return;  // the particle itself, a zero offset in the list
    const int o = A.orig_of[A.items[t]];
    if (k < cap) list[(size_t)p * cap + k] = sq_dist(dx, dy) <= I.h_eff_sq ? o : -1 - o;
    ++k;
  });
  counts[p] = k;
  path[p] = nl.cnt > nl.list_len ? 3 : (nl.base >= 0 ? (nl.pos ? 2 : 1) : 0);
}

__global__ void __launch_bounds__(256) k_pair_lists(LsaArgs A, int32_t *counts, int32_t *list, int32_t *path,
                                                    int cap) {
  extern __shared__ int32_t lds[];
  const int tid = threadIdx.x, B = blockDim.x;
  const int s = xcd_block() * B + tid;
  if (s >= A.n) return;
  const int32_t inf = A.sinfo[s];
  if (info_ptype(inf) != 0) return;
  for_instance(A.inst, inf,
               [&](const InstDev &I, int) { pair_list_particle(A, I, s, tid, lds, counts, list, path, cap); }, [] {});
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
// LDS of the neighbor lists: list_len 32-bit entries per thread plus the
// trash entry of the scan.
static size_t list_lds_bytes(int list_len, int block) { return (size_t)4 * (list_len + 1) * block; }

template <int CLS, bool DELTA>
static void csph_launch(hipStream_t s, const LsaArgs &a, int block, size_t lds, bool sum) {
  const dim3 g(xcd_grid(a.n, block)), b(block);
  if (sum)
    hipLaunchKernelGGL((k_csph<CLS, DELTA, true>), g, b, lds, s, a);
  else
    hipLaunchKernelGGL((k_csph<CLS, DELTA, false>), g, b, lds, s, a);
}

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

void launch_lsa(hipStream_t s, Device &d, int cur, StageMode mode, int guard, int keys, bool write_spd) {
  LsaArgs a;
  a.spd = write_spd ? d.spd : nullptr;
  a.mir = d.mir[cur];
  a.mout = d.mir[cur ^ 1];
  a.nbf = d.nbf[cur];
  a.nbf_out = d.nbf[cur ^ 1];
  a.sinfo = d.sinfo;
  a.items = d.items;
  a.orig_of = d.orig_of;
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
  a.fb_done = d.counters + kFbDone;
  a.mv = MoveOut{d.kslot, d.strays, d.nstray, d.counters, d.xbuild, guard, keys, d.retired, d.sinfo, d.n_inst, d.inst};
  a.n = d.n_total;
  a.kt = d.opt.kernel_type;
  a.list_len = d.list_len;
  a.staged = d.staged_scan;
  a.mode = mode;
  const size_t lds = list_lds_bytes(d.list_len, d.block);
  const int cls = d.opt.target_problem == 1 ? 0 : (d.opt.has_tem ? 2 : 1);
  const bool delta = d.opt.delta != 0;
  const int kt = d.opt.kernel_type;
  const int model = d.opt.sph_model;
  if (model <= 2) {  // classical SPH, FP64
    const bool sum = model == 1;
    if (cls == 0)
      csph_launch<0, false>(s, a, d.block, lds, sum);
    else if (cls == 1)
      delta ? csph_launch<1, true>(s, a, d.block, lds, sum) : csph_launch<1, false>(s, a, d.block, lds, sum);
    else
      delta ? csph_launch<2, true>(s, a, d.block, lds, sum) : csph_launch<2, false>(s, a, d.block, lds, sum);
    return;
  }
  if (model == 4) {  // q = 9, FP64, one thread per particle
    const dim3 g(xcd_grid(a.n, d.block)), b(d.block);
    if (cls == 0)
      hipLaunchKernelGGL((k_lsa<0, false, false, 0, 9>), g, b, lds, s, a);
    else if (cls == 1 && delta)
      hipLaunchKernelGGL((k_lsa<1, false, true, 0, 9>), g, b, lds, s, a);
    else if (cls == 1)
      hipLaunchKernelGGL((k_lsa<1, false, false, 0, 9>), g, b, lds, s, a);
    else if (delta)
      hipLaunchKernelGGL((k_lsa<2, false, true, 0, 9>), g, b, lds, s, a);
    else
      hipLaunchKernelGGL((k_lsa<2, false, false, 0, 9>), g, b, lds, s, a);
    return;
  }
  if (model == 5) {  // q = 14, FP64, one thread per particle (part of the sums in scratch;
                     // still faster than a lane group per particle)
    const dim3 g(xcd_grid(a.n, d.block)), b(d.block);
    if (cls == 0)
      hipLaunchKernelGGL((k_lsa<0, false, false, 0, 14>), g, b, lds, s, a);
    else if (cls == 1 && delta)
      hipLaunchKernelGGL((k_lsa<1, false, true, 0, 14>), g, b, lds, s, a);
    else if (cls == 1)
      hipLaunchKernelGGL((k_lsa<1, false, false, 0, 14>), g, b, lds, s, a);
    else if (delta)
      hipLaunchKernelGGL((k_lsa<2, false, true, 0, 14>), g, b, lds, s, a);
    else
      hipLaunchKernelGGL((k_lsa<2, false, false, 0, 14>), g, b, lds, s, a);
    return;
  }
  if (model == 6) {  // q = 20: a group of lanes per particle, FP64
    launch_lsa_wave(s, d, a, model, cls);
    return;
  }
  const bool mixed = d.opt.precision == LSSPH_MIXED;
  if (cls == 0)
    lsa_dispatch<0, false>(s, a, d.block, lds, mixed, kt);
  else if (cls == 1)
    delta ? lsa_dispatch<1, true>(s, a, d.block, lds, mixed, kt) : lsa_dispatch<1, false>(s, a, d.block, lds, mixed, kt);
  else
    delta ? lsa_dispatch<2, true>(s, a, d.block, lds, mixed, kt) : lsa_dispatch<2, false>(s, a, d.block, lds, mixed, kt);
}

void launch_pair_lists(hipStream_t s, Device &d, int cur, int32_t *counts, int32_t *list, int32_t *path,
                       int cap) {
  LsaArgs a{};
  a.mir = d.mir[cur];
  a.nbf = d.nbf[cur];
  a.sinfo = d.sinfo;
  a.items = d.items;
  a.orig_of = d.orig_of;
  a.cstart = d.cell_start;
  a.inst = d.inst;
  a.mv = MoveOut{d.kslot, d.strays, d.nstray, d.counters, d.xbuild, 0, 0, d.retired, d.sinfo, d.n_inst, d.inst};
  a.n = d.n_total;
  a.kt = d.opt.kernel_type;
  a.list_len = d.list_len;
  a.staged = d.staged_scan;
  hipLaunchKernelGGL(k_pair_lists, dim3(xcd_grid(a.n, d.block)), dim3(d.block), list_lds_bytes(d.list_len, d.block),
                     s, a, counts, list, path, cap);
}

void launch_umax(hipStream_t s, Device &d, int cur) {
  launch_fill_u64(s, d.umax, 0ull, d.n_inst);
  hipLaunchKernelGGL(k_umax, dim3(d.red_blocks, d.n_inst), dim3(256), 0, s, d.mir[cur], d.sinfo, d.cell_start,
                     d.inst, d.umax);
}

void launch_umax_spd(hipStream_t s, Device &d) {
  launch_fill_u64(s, d.umax, 0ull, d.n_inst);
  hipLaunchKernelGGL(k_umax_spd, dim3(d.red_blocks, d.n_inst), dim3(256), 0, s, d.spd, d.cell_start, d.inst,
                     d.umax);
}

void launch_dfmax(hipStream_t s, Device &d, int cur) {
  launch_fill_u64(s, d.dfmax, 0ull, d.n_inst);
  hipLaunchKernelGGL(k_dfmax, dim3(d.red_blocks, d.n_inst), dim3(256), 0, s, d.mir[cur], d.sinfo, d.items,
                     d.cell_start, d.f_prev, d.inst, d.dfmax);
}

template <bool T, bool MIXED>
static void pst_launch(hipStream_t s, const PstArgs &a, int block, size_t lds, int kt) {
  const dim3 g(xcd_grid(a.n, block)), b(block);
  if (kt == 3) {
    hipLaunchKernelGGL((k_pst<T, MIXED, 3, 5>), g, b, lds, s, a);
    if (MIXED) hipLaunchKernelGGL((k_pst_fallback<T, 3>), dim3(64), dim3(128), 0, s, a);
  } else {
    hipLaunchKernelGGL((k_pst<T, MIXED, 0, 5>), g, b, lds, s, a);
    if (MIXED) hipLaunchKernelGGL((k_pst_fallback<T, 0>), dim3(64), dim3(128), 0, s, a);
  }
}

void launch_pst(hipStream_t s, Device &d, int cur, int guard, int keys) {
  PstArgs a;
  a.mir = d.mir[cur];
  a.mout = d.mir[cur ^ 1];
  a.nbf = d.nbf[cur];
  a.nbf_out = d.nbf[cur ^ 1];
  a.sinfo = d.sinfo;
  a.items = d.items;
  a.orig_of = d.orig_of;
  a.cstart = d.cell_start;
  a.inst = d.inst;
  a.umax = d.umax;
  a.stats64 = d.stats64;
  a.fb_list = d.fb_list;
  a.fb_count = d.counters + kFbCount;
  a.fb_done = d.counters + kFbDone;
  a.mv = MoveOut{d.kslot, d.strays, d.nstray, d.counters, d.xbuild, guard, keys, d.retired, d.sinfo, d.n_inst, d.inst};
  a.n = d.n_total;
  a.kt = d.opt.kernel_type;
  a.list_len = d.list_len;
  a.staged = d.staged_scan;
  const size_t lds = list_lds_bytes(d.list_len, d.block);
  const int kt = d.opt.kernel_type;
  const int model = d.opt.sph_model;
  if (model <= 2) {  // q = 2 fit of the classical models, FP64
    const dim3 g(xcd_grid(a.n, d.block)), b(d.block);
    if (d.opt.has_tem)
      hipLaunchKernelGGL((k_pst<true, false, 0, 2>), g, b, lds, s, a);
    else
      hipLaunchKernelGGL((k_pst<false, false, 0, 2>), g, b, lds, s, a);
    return;
  }
  if (model == 4) {  // q = 9, FP64, one thread per particle
    const dim3 g(xcd_grid(a.n, d.block)), b(d.block);
    if (d.opt.has_tem)
      hipLaunchKernelGGL((k_pst<true, false, 0, 9>), g, b, lds, s, a);
    else
      hipLaunchKernelGGL((k_pst<false, false, 0, 9>), g, b, lds, s, a);
    return;
  }
  if (model == 5) {  // q = 14, FP64, one thread per particle
    const dim3 g(xcd_grid(a.n, d.block)), b(d.block);
    if (d.opt.has_tem)
      hipLaunchKernelGGL((k_pst<true, false, 0, 14>), g, b, lds, s, a);
    else
      hipLaunchKernelGGL((k_pst<false, false, 0, 14>), g, b, lds, s, a);
    return;
  }
  if (model == 6) {  // q = 20: a group of lanes per particle
    launch_pst_wave(s, d, a, model);
    return;
  }
  const bool mixed = d.opt.precision == LSSPH_MIXED;
  if (d.opt.has_tem)
    mixed ? pst_launch<true, true>(s, a, d.block, lds, kt) : pst_launch<true, false>(s, a, d.block, lds, kt);
  else
    mixed ? pst_launch<false, true>(s, a, d.block, lds, kt) : pst_launch<false, false>(s, a, d.block, lds, kt);
}

}  // namespace lssph
