// highorder.hip -- LS type A with the 5th-order Taylor basis (SPH_MODEL 6:
// q = 20; LSSPH_A_mod.f90:46-53, 148-172) and the particle shifting with the
// same basis (shift_interpolate_mod.f90:45-52, 153-185, 255-298), FP64.
//
// A 20 x 20 moment matrix (210 sums plus 20 per field) does not fit one
// thread's registers, so a group of lanes handles one particle.  Two designs
// share the neighbor search and give the same results bit for bit: the row
// groups below (20 lanes, 3 particles per wavefront, the default), and lane
// groups of GS lanes (32 or 64: 2 or 1 particles per wavefront, GS >= q):
//   1. the group tests the stencil's candidates GS at a time and compacts the
//      accepted ones into an LDS list with a ballot;
//   2. a rank sort puts the list into the reference's summation order
//      (position of the neighbor's reference cell in the 3x3 stencil of the
//      particle's reference cell, then upload index): at q = 20, whose
//      matrices reach cond(M) ~ 4e7, the summation order shows in the results;
//   3. in chunks of kChunk neighbors, lane k of the group forms the basis a,
//      VW * a and the field differences of neighbor k in LDS; then every lane
//      accumulates its share of the sums over the chunk, in list order;
//   4. the group factors the symmetrized matrix in LDS (LU with partial
//      pivoting, the operation order of lu_factor / lu_apply) and solves;
//   5. the group's first lane runs the epilogue of the q = 5 kernels (RHS,
//      stage update, cell bookkeeping).
// Each accumulator sums its neighbors in list order and the factorization's
// operations do not depend on GS, so every group size gives the same
// results bit for bit; GS only trades lanes against particles in flight.
// Groups never wait for each other (no block barrier), so a group whose slot
// is a wall or past the end simply returns.  (The templates also take
// q = 14, but SPH_MODEL 5 runs 4x faster on the per-thread kernels of
// operators.hip, even with part of its sums in scratch: 7.9 against 34 ms per
// 1024^2 step.)
#include "ls_common.hpp"

namespace lssph {

constexpr int kChunk = 16;   // neighbors staged in LDS per round
constexpr int kWavesPB = 2;  // wavefronts per block

// Lanes of one particle's group: GS consecutive lanes of a wavefront.
template <int GS>
struct Group {
  static constexpr int size = GS;
  static constexpr int G = 64 / GS;  // particles per wavefront
  int lane;  // lane in the wavefront
  int g;     // group in the wavefront
  int gl;    // lane in the group
  unsigned long long mask;  // the group's lanes
  __device__ __forceinline__ explicit Group(int l) : lane(l), g(l / GS), gl(l % GS) {
    mask = (GS == 64 ? ~0ull : ((1ull << GS) - 1ull)) << (g * GS);
  }
  // maximum over the group's lanes of (v, index), ties to the lower index
  __device__ __forceinline__ void argmax(double &v, int &idx) const {
#pragma unroll
    for (int off = 1; off < GS; off <<= 1) {
      const double v2 = __shfl_xor(v, off, GS);
      const int i2 = __shfl_xor(idx, off, GS);
      const bool take = (v2 > v) || (v2 == v && i2 < idx);
      v = take ? v2 : v;
      idx = take ? i2 : idx;
    }
  }
  // the value of the group's lane k
  template <typename T>
  __device__ __forceinline__ T from(T v, int k) const {
    return __shfl(v, g * GS + k, 64);
  }
};


// Kernel W(r) and dW/dr written term by term as kernel_functions_mod.f90
// forms them (q = r / h, powers as repeated products, normalisation applied
// last): with cond(M_20) up to 4e7, a last-bit difference in W shows.
__device__ __forceinline__ double ipow_d(double t, int n) {
  double r = t;
  for (int i = 1; i < n; ++i) r *= t;
  return r;
}

__device__ double ref_w(int kt, double r, double h) {
  const double q = r / h;
  const double pi = kPi;
  double w;
  switch (kt) {
    case 1:  // :33-41
      w = (0.0 <= q && q <= 1.0) ? ipow_d(2.0 - q, 3) - 4.0 * ipow_d(1.0 - q, 3)
                                 : ((1.0 < q && q <= 2.0) ? ipow_d(2.0 - q, 3) : 0.0);
      return 5.0 / (14.0 * pi * (h * h)) * w;
    case 2:  // :43-53
      if (0.0 <= q && q <= 1.0)
        w = ipow_d(3.0 - q, 5) - 6.0 * ipow_d(2.0 - q, 5) + 15.0 * ipow_d(1.0 - q, 5);
      else if (1.0 < q && q <= 2.0)
        w = ipow_d(3.0 - q, 5) - 6.0 * ipow_d(2.0 - q, 5);
      else
        w = (2.0 < q && q <= 3.0) ? ipow_d(3.0 - q, 5) : 0.0;
      return 7.0 / (478.0 * pi * (h * h)) * w;
    case 3:  // :55-61
      w = (0.0 <= q && q <= 2.0) ? ipow_d(1.0 - 0.5 * q, 4) * (1.0 + 2.0 * q) : 0.0;
      return 7.0 / (4.0 * pi * (h * h)) * w;
    case 4:  // :63-69
      w = (0.0 <= q && q <= 2.0) ? ipow_d(1.0 - 0.5 * q, 6) * (1.0 + 3.0 * q + 35.0 * (q * q) / 12.0) : 0.0;
      return 9.0 / (4.0 * pi * (h * h)) * w;
    default:  // :71-78
      w = (0.0 <= q && q <= 2.0)
              ? ipow_d(1.0 - 0.5 * q, 8) * (1.0 + 4.0 * q + 25.0 * (q * q) / 4.0 + 4.0 * (q * q * q))
              : 0.0;
      return 39.0 / (14.0 * pi * (h * h)) * w;
  }
}

__device__ double ref_dw(int kt, double r, double h) {
  const double q = r / h;
  const double pi = kPi;
  double d;
  switch (kt) {
    case 1:  // :94-102
      d = (0.0 <= q && q <= 1.0) ? ipow_d(2.0 - q, 2) - 4.0 * ipow_d(1.0 - q, 2)
                                 : ((1.0 < q && q <= 2.0) ? ipow_d(2.0 - q, 2) : 0.0);
      return 5.0 / (14.0 * pi * (h * h)) * d * (-3.0 / h);
    case 2:  // :104-114
      if (0.0 <= q && q <= 1.0)
        d = ipow_d(3.0 - q, 4) - 6.0 * ipow_d(2.0 - q, 4) + 15.0 * ipow_d(1.0 - q, 4);
      else if (1.0 < q && q <= 2.0)
        d = ipow_d(3.0 - q, 4) - 6.0 * ipow_d(2.0 - q, 4);
      else
        d = (2.0 < q && q <= 3.0) ? ipow_d(3.0 - q, 4) : 0.0;
      return 7.0 / (478.0 * pi * (h * h)) * d * (-5.0 / h);
    case 3:  // :116-123
      d = (0.0 <= q && q <= 2.0)
              ? -2.0 * ipow_d(1.0 - 0.5 * q, 3) * (1.0 + 2.0 * q) + 2.0 * ipow_d(1.0 - 0.5 * q, 4)
              : 0.0;
      return 7.0 / (4.0 * pi * (h * h)) * d / h;
    case 4:  // :125-132
      d = (0.0 <= q && q <= 2.0) ? -3.0 * ipow_d(1.0 - 0.5 * q, 5) * (1.0 + 3.0 * q + 35.0 * (q * q) / 12.0) +
                                       1.0 * ipow_d(1.0 - 0.5 * q, 6) * (3.0 + 35.0 * q / 6.0)
                                 : 0.0;
      return 9.0 / (4.0 * pi * (h * h)) * d / h;
    default:  // :134-143
      d = (0.0 <= q && q <= 2.0)
              ? -4.0 * ipow_d(1.0 - 0.5 * q, 7) * (1.0 + 4.0 * q + 25.0 * (q * q) / 4.0 + 4.0 * (q * q * q)) +
                    1.0 * ipow_d(1.0 - 0.5 * q, 8) * (4.0 + 25.0 * q / 2.0 + 12.0 * (q * q))
              : 0.0;
      return 39.0 / (14.0 * pi * (h * h)) * d / h;
  }
}

// Taylor basis of q terms with ih[k] = 1/h^(k+1) (LSSPH_A_mod.f90:140-170,
// shift_interpolate_mod.f90:153-185 and 259-297).
template <int Q>
__device__ __forceinline__ void basis_hq(double x, double y, const double (&ih)[5], double (&a)[Q]) {
  a[0] = x * ih[0];
  a[1] = y * ih[0];
  a[2] = 0.5 * (x * x) * ih[1];
  a[3] = 0.5 * (y * y) * ih[1];
  a[4] = x * y * ih[1];
  a[5] = 1.0 / 6.0 * (x * x * x) * ih[2];
  a[6] = 1.0 / 6.0 * (y * y * y) * ih[2];
  a[7] = 0.5 * (x * x) * y * ih[2];
  a[8] = 0.5 * (y * y) * x * ih[2];
  if constexpr (Q >= 14) {
    a[9] = 1.0 / 24.0 * (x * x * x * x) * ih[3];
    a[10] = 1.0 / 24.0 * (y * y * y * y) * ih[3];
    a[11] = 1.0 / 6.0 * (x * x * x) * y * ih[3];
    a[12] = 1.0 / 6.0 * (y * y * y) * x * ih[3];
    a[13] = 1.0 / 4.0 * (x * x) * (y * y) * ih[3];
  }
  if constexpr (Q >= 20) {
    a[14] = 1.0 / 120.0 * (x * x * x * x * x) * ih[4];
    a[15] = 1.0 / 120.0 * (y * y * y * y * y) * ih[4];
    a[16] = 1.0 / 24.0 * (x * x * x * x) * y * ih[4];
    a[17] = 1.0 / 24.0 * (y * y * y * y) * x * ih[4];
    a[18] = 1.0 / 12.0 * (x * x * x) * (y * y) * ih[4];
    a[19] = 1.0 / 12.0 * (y * y * y) * (x * x) * ih[4];
  }
}

// LDS of one particle's group.  Row k of `rows` holds neighbor k of the
// chunk: a[0..Q), VW * a[0..Q), the field differences [0..NF) and the two
// shift terms of the shifting kernel.  The factorization reuses the same
// space.  About 7.5 KB for q = 20.
template <int Q, int NF>
struct WaveLds {
  static constexpr int RS = 2 * Q + NF + 2;
  int32_t list[kWaveList];    // accepted candidates, stencil order
  int32_t sorted[kWaveList];  // the same in summation order
  uint32_t key[kWaveList];    // stencil position << 27 | upload index (< 2^27)
  union {
    double rows[kChunk][RS];
    struct {
      double m[Q][Q];
      double b[Q][NF];
    } s;
  } u;
};

// Neighbors of the particle at slot s (record me) within h_eff, in the
// reference's summation order, into sorted[] (list[] and key[] are work
// space).  Returns the count; a count above kWaveList is flagged and only the
// first kWaveList are kept.  GR: the lanes of the particle's group (size,
// mask, lane, gl).
template <class GR>
__device__ int group_neighbors(const Rec *__restrict__ mir, const int32_t *__restrict__ cstart,
                               const int32_t *__restrict__ sinfo, const int32_t *__restrict__ items,
                               const int32_t *__restrict__ orig_of, int32_t *counters, const InstDev &I, int s,
                               const Rec &me, int32_t *list, uint32_t *key, int32_t *sorted, const GR &G) {
  constexpr int GS = GR::size;
  const int lc = local_cell(I, me.x, me.y);
  const unsigned long long below = G.mask & ((1ull << G.lane) - 1ull);
  int cnt = 0;
  visit_cells(I, lc, cstart, [&](int t0, int t1, int, int) {
    for (int base = t0; base < t1; base += GS) {
      const int t = base + G.gl;
      bool ok = false;
      int code = 0;
      if (t < t1 && t != s) {
        const double2 p = load_xy(mir, t);
        const double sx = I.periodic ? image_shift(p.x, me.x, I.len_x) : 0.0;
        const double sy = I.periodic ? image_shift(p.y, me.y, I.len_y) : 0.0;
        ok = sq_dist((p.x + sx) - me.x, (p.y + sy) - me.y) <= I.h_eff_sq;
        code = image_code(sx, sy);
      }
      const unsigned long long m = __ballot(ok) & G.mask;
      const int pos = cnt + __popcll(m & below);
      if (ok && pos < kWaveList) list[pos] = t | (code << kSlotBits);
      cnt += __popcll(m);
    }
  });
  if (cnt > kWaveList) {
    if (G.gl == 0) atomicAdd(&counters[kWaveListOverflow], 1);
    cnt = kWaveList;
  }
  // sort keys: stencil position of the reference cells, then upload index
  int rcx, rcy;
  ref_cell_xy(I, me.x, me.y, false, rcx, rcy);
  for (int e = G.gl; e < cnt; e += GS) {
    const int t = list[e] & kSlotMask;
    const double2 p = load_xy(mir, t);
    int jx, jy;
    ref_cell_xy(I, p.x, p.y, info_ptype(sinfo[t]) != 0, jx, jy);
    int dx = jx - rcx, dy = jy - rcy;
    if (I.periodic) {
      dx = dx > 1 ? dx - I.rnx : (dx < -1 ? dx + I.rnx : dx);
      dy = dy > 1 ? dy - I.rny : (dy < -1 ? dy + I.rny : dy);
    }
    key[e] = ((uint32_t)((dy + 1) * 3 + (dx + 1)) << kSlotBits) | (uint32_t)orig_of[items[t]];
  }
  wave_sync();
  // rank sort (the keys are distinct)
  for (int e = G.gl; e < cnt; e += GS) {
    const uint32_t k = key[e];
    int rank = 0;
    for (int j = 0; j < cnt; ++j) rank += (key[j] < k) ? 1 : 0;
    sorted[rank] = list[e];
  }
  wave_sync();
  return cnt;
}

// Per-lane share of the LS sums: entry e < Q(Q+1)/2 is M(r, i), r <= i,
// accumulated as (VW a_i) * a_r (LSSPH_A_mod.f90:185-188); the others are
// b(i, f) += (VW a_i) * df_f (:178-183).  Lane gl of the group owns entries
// gl, gl + GS, ...; o1/o2 are the row offsets of the two factors.
template <int Q, int NF, int GS>
struct LaneSums {
  static constexpr int NM = Q * (Q + 1) / 2;
  static constexpr int NE = NM + Q * NF;
  static constexpr int PER = (NE + GS - 1) / GS;
  double acc[PER];
  int o1[PER], o2[PER];
  __device__ __forceinline__ void init(int gl) {
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      acc[m] = 0.0;
      const int e = gl + GS * m;
      int i = 0, r = 0, f = -1;
      if (e < NM) {
        while ((i + 1) * (i + 2) / 2 <= e) ++i;
        r = e - i * (i + 1) / 2;
      } else if (e < NE) {
        f = (e - NM) / Q;
        i = (e - NM) % Q;
      }
      o1[m] = Q + i;
      o2[m] = f < 0 ? r : 2 * Q + f;
    }
  }
  template <class WL>
  __device__ __forceinline__ void add_chunk(const WL &W, int nc, int gl) {
    for (int k = 0; k < nc; ++k) {
      const double *row = W.u.rows[k];
#pragma unroll
      for (int m = 0; m < PER; ++m)
        if (gl + GS * m < NE) acc[m] += row[o1[m]] * row[o2[m]];
    }
  }
  // symmetrized matrix and right-hand sides into W.u.s (after a wave_sync)
  template <class WL>
  __device__ __forceinline__ void store(WL &W, int gl) const {
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int e = gl + GS * m;
      if (e < NM) {
        const int i = o1[m] - Q, r = o2[m];
        W.u.s.m[r][i] = acc[m];
        W.u.s.m[i][r] = acc[m];
      } else if (e < NE) {
        W.u.s.b[o1[m] - Q][o2[m] - 2 * Q] = acc[m];
      }
    }
  }
};

// DGESV on W.u.s by one group: LU with partial pivoting in the order of
// lu_factor (first row of the largest |a| as pivot; a zero pivot sets info
// and skips the scaling), then -- when no pivot is zero -- the row
// interchanges, forward and back substitution of lu_apply.  Returns info.
template <int Q, int NF, int GS, class WL>
__device__ int group_dgesv(WL &W, const Group<GS> &G) {
  static_assert(GS >= Q, "a group holds one row of the matrix per lane");
  auto &m = W.u.s.m;
  auto &b = W.u.s.b;
  const int gl = G.gl;
  int piv[Q];
  int info = 0;
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    double v = (gl >= j && gl < Q) ? fabs(m[gl < Q ? gl : 0][j]) : -1.0;
    int idx = gl;
    G.argmax(v, idx);
    int p = G.from(idx, 0);
    p = (p >= j && p < Q) ? p : j;  // only a column of NaNs leaves no pivot row
    piv[j] = p;
    if (p != j && gl < Q) {
      const double t = m[j][gl];
      m[j][gl] = m[p][gl];
      m[p][gl] = t;
    }
    wave_sync();
    const double pv = m[j][j];
    if (pv == 0.0) {
      info = info ? info : j + 1;
    } else if (gl > j && gl < Q) {
      m[gl][j] = m[gl][j] / pv;
    }
    wave_sync();
    const int nn = Q - 1 - j;
    for (int e = gl; e < nn * nn; e += GS) {
      const int i = j + 1 + e / nn, c = j + 1 + e % nn;
      m[i][c] -= m[i][j] * m[j][c];
    }
    wave_sync();
  }
  if (info) return info;
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const int p = piv[j];
    if (p != j && gl < NF) {
      const double t = b[j][gl];
      b[j][gl] = b[p][gl];
      b[p][gl] = t;
    }
  }
  wave_sync();
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const int rows = (Q - 1 - j) * NF;
    for (int e = gl; e < rows; e += GS) {
      const int i = j + 1 + e / NF, f = e % NF;
      b[i][f] -= b[j][f] * m[i][j];
    }
    wave_sync();
  }
#pragma unroll
  for (int j = Q - 1; j >= 0; --j) {
    if (gl < NF) b[j][gl] = b[j][gl] / m[j][j];
    wave_sync();
    for (int e = gl; e < j * NF; e += GS) {
      const int i = e / NF, f = e % NF;
      b[i][f] -= b[j][f] * m[i][j];
    }
    wave_sync();
  }
  return 0;
}

__device__ __forceinline__ void inv_h_powers(const InstDev &I, double (&ih)[5]) {
  ih[0] = I.inv_h;
  ih[1] = I.inv_h2;  // (1/h)/h as the reference's inv_h(2)
#pragma unroll
  for (int k = 2; k < 5; ++k) ih[k] = ih[k - 1] / I.h;
}

// Terms of list entry e in the LS type-A sums of the particle me: the basis a,
// V_j W and the field differences (LSSPH_A_mod.f90:130-183).
template <int Q, int CLS>
__device__ __forceinline__ void lsa_terms(const LsaArgs &A, const InstDev &I, const Rec &me, int e, int kt,
                                          const double (&ih)[5], double (&a)[Q], double &vw,
                                          double (&df)[Layout<CLS, true>::NF]) {
  const int t = e & kSlotMask, code = e >> kSlotBits;
  const Rec nb = A.mir[t];
  const double dx = (nb.x + shift_of(code & 3, I.len_x)) - me.x;
  const double dy = (nb.y + shift_of(code >> 2, I.len_y)) - me.y;
  const double r = sqrt(sq_dist(dx, dy));
  vw = nb.vol * ref_w(kt, r, I.h);
  basis_hq<Q>(dx, dy, ih, a);
  field_diffs<CLS, true>(nb, me, df);
}

// Terms of list entry e in the shifting of the particle me: the basis a,
// V_j W, the differences of u, v, rho (, tem) and the neighbor's two shift
// terms (shift_interpolate_mod.f90:130-201).
template <int Q, int NF>
__device__ __forceinline__ void pst_terms(const PstArgs &A, const InstDev &I, const Rec &me, int e, int kt,
                                          const double (&ih)[5], double (&a)[Q], double &vw, double (&df)[NF],
                                          double &shx, double &shy) {
  const int t = e & kSlotMask, code = e >> kSlotBits;
  const Rec nb = A.mir[t];
  const double dx = (nb.x + shift_of(code & 3, I.len_x)) - me.x;
  const double dy = (nb.y + shift_of(code >> 2, I.len_y)) - me.y;
  const double r = sqrt(sq_dist(dx, dy));
  const double wv = ref_w(kt, r, I.h);
  const double dwdr = ref_dw(kt, r, I.h);
  const double vj = nb.vol;
  vw = vj * wv;
  const double dwx = dwdr * (-dx) / r * vj;  // :142-143
  const double dwy = dwdr * (-dy) / r * vj;
  const double wr = wv / I.w_ave;
  const double fac = 1.0 + 0.2 * ((wr * wr) * (wr * wr));
  basis_hq<Q>(dx, dy, ih, a);
  df[0] = nb.u - me.u;
  df[1] = nb.v - me.v;
  df[2] = nb.rho - me.rho;
  if constexpr (NF == 4) df[3] = nb.tem - me.tem;
  shx = fac * dwx;
  shy = fac * dwy;
}

// ---------------------------------------------------------------------------
// LS type A (LSSPH_A_mod.f90:26-271) with q = 14, 20 and the stage update.
// The field columns are those of the q = 5 kernels with rho always carried.
// ---------------------------------------------------------------------------
template <int Q, int CLS, int GS>
__global__ void __launch_bounds__(64 * kWavesPB) k_lsa_wave(LsaArgs A) {
  using L = Layout<CLS, true>;
  constexpr int NF = L::NF;
  using WL = WaveLds<Q, NF>;
  constexpr int NG = 64 / GS;
  __shared__ WL lds_all[kWavesPB * NG];
  const Group<GS> G(threadIdx.x & 63);
  const int w = threadIdx.x >> 6;
  const int s = (blockIdx.x * kWavesPB + w) * NG + G.g;
  if (s >= A.n) return;
  WL &W = lds_all[w * NG + G.g];
  const int32_t inf = A.sinfo[s];
  const InstDev &I = A.inst[info_inst(inf)];
  // walls, and the particles of retired instances, move on with the records
  if (info_ptype(inf) != 0 || I.retired) {
    if (G.gl == 0 && A.mode != kRhsOnly) A.mout[s] = A.mir[s];
    if (G.gl == 0 && A.spd) A.spd[s] = 0.0;
    return;
  }
  const Rec me = A.mir[s];
  const int n = group_neighbors(A.mir, A.cstart, A.sinfo, A.items, A.orig_of, A.counters, I, s, me, W.list, W.key,
                                W.sorted, G);
  double ih[5];
  inv_h_powers(I, ih);
  const int kt = A.kt;
  LaneSums<Q, NF, GS> S;
  S.init(G.gl);
  for (int c0 = 0; c0 < n; c0 += kChunk) {
    const int nc = min(kChunk, n - c0);
    if (G.gl < nc) {
      double a[Q], vw, df[NF];
      lsa_terms<Q, CLS>(A, I, me, W.sorted[c0 + G.gl], kt, ih, a, vw, df);
      double *row = W.u.rows[G.gl];
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        row[i] = a[i];
        row[Q + i] = vw * a[i];
      }
#pragma unroll
      for (int f = 0; f < NF; ++f) row[2 * Q + f] = df[f];
    }
    wave_sync();
    S.add_chunk(W, nc, G.gl);
    wave_sync();
  }
  S.store(W, G.gl);
  wave_sync();
  double d[NF][5];
  if (n >= Q) {  // LSSPH_A_mod.f90:208 (num_neighbors >= q)
    group_dgesv<Q, NF>(W, G);
    double X[NF][5];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int i = 0; i < 5; ++i) X[f][i] = W.u.s.b[i][f];
    scale_derivs<NF>(X, I, d);
  } else {
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int i = 0; i < 5; ++i) d[f][i] = 0.0;
  }
  if (G.gl == 0) finish<CLS, true>(A, I, A.items[s], s, me, d);
}

// ---------------------------------------------------------------------------
// Particle shifting (shift_interpolate_mod.f90:23-300) with q = 14, 20.
// ---------------------------------------------------------------------------
template <int Q, bool T, int GS>
__global__ void __launch_bounds__(64 * kWavesPB) k_pst_wave(PstArgs A) {
  constexpr int NF = T ? 4 : 3;  // u, v, rho (, tem)
  using WL = WaveLds<Q, NF>;
  constexpr int NG = 64 / GS;
  __shared__ WL lds_all[kWavesPB * NG];
  const Group<GS> G(threadIdx.x & 63);
  const int w = threadIdx.x >> 6;
  const int s = (blockIdx.x * kWavesPB + w) * NG + G.g;
  if (s >= A.n) return;
  WL &W = lds_all[w * NG + G.g];
  const int32_t inf = A.sinfo[s];
  const int ii = info_inst(inf);
  const InstDev &I = A.inst[ii];
  if (info_ptype(inf) != 0 || I.retired) {
    if (G.gl == 0) A.mout[s] = A.mir[s];  // walls move on with the records
    return;
  }
  const Rec me = A.mir[s];
  const int n = group_neighbors(A.mir, A.cstart, A.sinfo, A.items, A.orig_of, A.mv.counters, I, s, me, W.list,
                                W.key, W.sorted, G);
  double ih[5];
  inv_h_powers(I, ih);
  const int kt = A.kt;
  LaneSums<Q, NF, GS> S;
  S.init(G.gl);
  double sx = 0.0, sy = 0.0;
  for (int c0 = 0; c0 < n; c0 += kChunk) {
    const int nc = min(kChunk, n - c0);
    if (G.gl < nc) {
      double a[Q], vw, df[NF], shx, shy;
      pst_terms<Q, NF>(A, I, me, W.sorted[c0 + G.gl], kt, ih, a, vw, df, shx, shy);
      double *row = W.u.rows[G.gl];
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        row[i] = a[i];
        row[Q + i] = vw * a[i];
      }
#pragma unroll
      for (int f = 0; f < NF; ++f) row[2 * Q + f] = df[f];
      row[2 * Q + NF] = shx;
      row[2 * Q + NF + 1] = shy;
    }
    wave_sync();
    for (int k = 0; k < nc; ++k) {  // the shift sums in list order (:145-146)
      sx = sx + W.u.rows[k][2 * Q + NF];
      sy = sy + W.u.rows[k][2 * Q + NF + 1];
    }
    S.add_chunk(W, nc, G.gl);
    wave_sync();
  }
  S.store(W, G.gl);
  wave_sync();
  // shift vector and its clamp (:217-225)
  const double umax = __longlong_as_double((long long)A.umax[ii]);
  const double coe = I.pst_c * umax * I.dt * I.h;
  double dxs = -coe * sx, dys = -coe * sy;
  const double rs = sqrt(dxs * dxs + dys * dys);
  if (rs > 0.2 * I.h) {
    dxs = (0.2 * I.h) * dxs / rs;
    dys = (0.2 * I.h) * dys / rs;
  }
  bool solved = false;
  if (n >= Q) {  // :240-247
    group_dgesv<Q, NF>(W, G);
    solved = true;
  }
  if (G.gl != 0) return;
  double a[Q];
  basis_hq<Q>(dxs, dys, ih, a);
  double acc[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    double t = 0.0;
    if (solved) {
#pragma unroll
      for (int i = 0; i < Q; ++i) t += a[i] * W.u.s.b[i][f];
    }
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
// Row groups: Q lanes per particle, 64 / Q particles per wavefront (3 at
// q = 20); lane i of a group owns row i of its particle's moment matrix.
//   1. neighbor search and summation order as above (group_neighbors);
//   2. in chunks of Q neighbors, lane k forms a, V_j W and the field
//      differences of neighbor k in LDS; then lane i accumulates column i of
//      the upper triangle, M(r, i) += (V_j W a_i) a_r for r <= i, and
//      b(i, f) += (V_j W a_i) df_f, in list order: the operands and order of
//      the lane-group kernels above, so the sums agree bit for bit;
//   3. the group factors the symmetric matrix with partial pivoting without
//      moving rows: each lane keeps its row in registers and the row's
//      position; LDS holds a copy of every row by position and the pivot
//      column's candidates, from which every lane takes the pivot (the first
//      position of the largest |a|) and the pivot row; each elimination step
//      also updates the lane's right-hand sides -- the forward substitution
//      of lu_apply, in its order -- and writes the row back;
//   4. lane f of the group runs the back substitution of field f over U in
//      LDS.
// Every operation of group_dgesv is done once, on the same operands and in
// the same order, so both kernels give the same results bit for bit.
// ---------------------------------------------------------------------------
template <int Q>
struct RowGroup {
  static constexpr int size = Q;
  static constexpr int G = 64 / Q;  // particles per wavefront
  int lane;  // lane in the wavefront
  int g;     // group in the wavefront (G: an idle lane)
  int gl;    // lane in the group: the row it owns
  unsigned long long mask;
  __device__ __forceinline__ explicit RowGroup(int l) : lane(l), g(l / Q), gl(l % Q) {
    mask = ((1ull << Q) - 1ull) << (g * Q);
  }
};

// LDS of one particle's row group (4.8 KB at q = 20).
template <int Q, int NF, int NX>
struct alignas(16) RowLds {
  static constexpr int RS = (Q + 1 + NF + NX + 1) & ~1;  // even: every row 16-B aligned
  static constexpr int VW = Q, DF = Q + 1, XT = Q + 1 + NF;  // row offsets after a[0..Q)
  int32_t sorted[kWaveList];
  union {
    struct {
      int32_t list[kWaveList];
      uint32_t key[kWaveList];
    } scan;
    double rows[Q][RS];  // chunk: a, V_j W, df, extra terms of neighbor k
    struct {
      double r[Q][RS];   // the row at position p: r[p][0..Q), right-hand sides r[p][Q..Q + NF)
      double cand[Q];    // |r[p][j]| of the rows not yet pivots
    } lu;
  } u;
};

// (v, key) over the group's lanes: the largest v, ties to the smaller key;
// every lane of the group gets the result.
template <int Q>
__device__ __forceinline__ void row_argmax(double &v, int &key, const RowGroup<Q> &G) {
#pragma unroll
  for (int off = 1; off < Q; off <<= 1) {
    const double v2 = __shfl(v, G.lane + off, 64);
    const int k2 = __shfl(key, G.lane + off, 64);
    const bool take = (G.gl + off < Q) && ((v2 > v) || (v2 == v && k2 < key));
    v = take ? v2 : v;
    key = take ? k2 : key;
  }
  v = __shfl(v, G.g * Q, 64);
  key = __shfl(key, G.g * Q, 64);
}

// Per-lane sums of a row group (step 2): acc[r] = M(r, i) for r <= i (the
// entries above i are not used), bacc[f] = b(i, f).
template <int Q, int NF>
struct RowSums {
  double acc[Q];
  double bacc[NF];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int r = 0; r < Q; ++r) acc[r] = 0.0;
#pragma unroll
    for (int f = 0; f < NF; ++f) bacc[f] = 0.0;
  }
  // Neighbor k's operands are read into registers together before they are
  // used (left to itself the scheduler reuses two registers for the reads
  // and waits on each).
  template <class RL>
  __device__ __forceinline__ void add_chunk(const RL &W, int nc, int i) {
    for (int k = 0; k < nc; ++k) {
      const double *row = W.u.rows[k];
      double a[Q], g[1 + NF];
#pragma unroll
      for (int r = 0; r < Q; ++r) a[r] = row[r];
#pragma unroll
      for (int f = 0; f <= NF; ++f) g[f] = row[RL::VW + f];
      const double ai = row[i];
      __builtin_amdgcn_sched_barrier(0);
      const double vwa = g[0] * ai;  // V_j W a_i
#pragma unroll
      for (int r = 0; r < Q; ++r) acc[r] += vwa * a[r];
#pragma unroll
      for (int f = 0; f < NF; ++f) bacc[f] += vwa * g[1 + f];
    }
  }
};

// f(std::integral_constant<int, J>) for J = B .. E - 1, in order.
template <int B, int E>
struct StaticFor {
  template <class F>
  static __device__ __forceinline__ void run(F &&f) {
    if constexpr (B < E) {
      f(std::integral_constant<int, B>{});
      StaticFor<B + 1, E>::run(f);
    }
  }
};

// First position of the largest value among c[LO..LO + N) (ties and NaNs keep
// the earlier position), as a tree over the two halves.
template <int LO, int N>
struct FirstMax {
  template <int M>
  static __device__ __forceinline__ void run(const double (&c)[M], double &v, int &p) {
    if constexpr (N == 1) {
      v = c[LO];
      p = LO;
    } else {
      constexpr int H = (N + 1) / 2;
      double v1, v2;
      int p1, p2;
      FirstMax<LO, H>::run(c, v1, p1);
      FirstMax<LO + H, N - H>::run(c, v2, p2);
      const bool take = v2 > v1;
      v = take ? v2 : v1;
      p = take ? p2 : p1;
    }
  }
};

// DGESV on the sums of a row group (steps 3 and 4).  The group's rows are
// kept by position in LDS (R[p] = the row at position p and its right-hand
// sides, cand[p] = |R[p][j]| for the next pivot column): every lane reads
// the candidates and the pivot row from there, and the lanes of the rows
// that are not yet pivots write theirs back after each elimination step.
// On return R[k][Q + f] holds the solution of field f -- or, when a pivot
// is zero, the unsolved right-hand sides, as DGESV leaves them.
template <int Q, int NF, class RL>
__device__ __forceinline__ void row_dgesv(RL &W, const RowSums<Q, NF> &S, const RowGroup<Q> &G) {
  auto &R = W.u.lu.r;
  auto &cand = W.u.lu.cand;
  const int i = G.gl;
  // M(r, i), r <= i, is acc[r] of lane i: row i, column r and row r, column i
#pragma unroll
  for (int r = 0; r < Q; ++r)
    if (r <= i) {
      R[i][r] = S.acc[r];
      R[r][i] = S.acc[r];
    }
#pragma unroll
  for (int f = 0; f < NF; ++f) R[i][Q + f] = S.bacc[f];
  wave_sync();
  double row[Q], b[NF];
#pragma unroll
  for (int r = 0; r < Q; ++r) row[r] = R[i][r];
#pragma unroll
  for (int f = 0; f < NF; ++f) b[f] = S.bacc[f];
  cand[i] = fabs(row[0]);
  wave_sync();
  int pos = i;        // the position of this lane's row
  bool done = false;  // the row was a pivot row
  bool singular = false;
  StaticFor<0, Q>::run([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    // the pivot: the first position >= j of the largest |a(., j)|
    double c[Q - j];
#pragma unroll
    for (int k = 0; k < Q - j; ++k) c[k] = cand[j + k];
    double cmax;
    int bk;
    FirstMax<0, Q - j>::run(c, cmax, bk);
    const int bp = j + bk;
    double u[Q - j], ub[NF];
#pragma unroll
    for (int k = 0; k < Q - j; ++k) u[k] = R[bp][j + k];
#pragma unroll
    for (int f = 0; f < NF; ++f) ub[f] = R[bp][Q + f];
    wave_sync();  // the reads of R and cand come before this step's writes
    const double pv = u[0];
    singular = singular || pv == 0.0;
    const bool active = !done;  // the row is at a position >= j
    const bool piv = pos == bp;
    pos = piv ? j : (pos == j ? bp : pos);
    done = done || piv;
    // the rows below the pivot: multiplier, trailing update, forward
    // substitution (a zero pivot skips the division); the others add zeros
    const double lq = pv != 0.0 ? row[j] / pv : row[j];
    const double l = done ? 0.0 : lq;
    row[j] = done ? row[j] : lq;
#pragma unroll
    for (int k = 1; k < Q - j; ++k) row[j + k] -= l * u[k];
#pragma unroll
    for (int f = 0; f < NF; ++f) b[f] -= ub[f] * l;
    if (active) {
#pragma unroll
      for (int k = 0; k < Q - j; ++k) R[pos][j + k] = row[j + k];
#pragma unroll
      for (int f = 0; f < NF; ++f) R[pos][Q + f] = b[f];
      if constexpr (j + 1 < Q)
        if (!done) cand[pos] = fabs(row[j + 1]);
    }
    wave_sync();
  });
  // back substitution, one field per lane (lu_apply)
  if (!singular) {
    if (i < NF) {
      double x[Q];
#pragma unroll
      for (int p = 0; p < Q; ++p) x[p] = R[p][Q + i];
#pragma unroll
      for (int j = Q - 1; j >= 0; --j) {
        x[j] = x[j] / R[j][j];
#pragma unroll
        for (int p = 0; p < j; ++p) x[p] -= x[j] * R[p][j];
      }
#pragma unroll
      for (int p = 0; p < Q; ++p) R[p][Q + i] = x[p];
    }
  } else {
#pragma unroll
    for (int f = 0; f < NF; ++f) R[i][Q + f] = S.bacc[f];
  }
  wave_sync();
}

// LS type A with q = 20 on row groups (see above) and the stage update.
template <int Q, int CLS>
__global__ void __launch_bounds__(64) k_lsa_rows(LsaArgs A) {
  using L = Layout<CLS, true>;
  constexpr int NF = L::NF;
  using RL = RowLds<Q, NF, 0>;
  constexpr int NG = RowGroup<Q>::G;
  __shared__ RL lds_all[NG];
  const RowGroup<Q> G(threadIdx.x);
  if (G.g >= NG) return;
  const int s = blockIdx.x * NG + G.g;
  if (s >= A.n) return;
  RL &W = lds_all[G.g];
  const int32_t inf = A.sinfo[s];
  const InstDev &I = A.inst[info_inst(inf)];
  if (info_ptype(inf) != 0 || I.retired) {
    if (G.gl == 0 && A.mode != kRhsOnly) A.mout[s] = A.mir[s];
    if (G.gl == 0 && A.spd) A.spd[s] = 0.0;
    return;
  }
  const Rec me = A.mir[s];
  const int n = group_neighbors(A.mir, A.cstart, A.sinfo, A.items, A.orig_of, A.counters, I, s, me,
                                W.u.scan.list, W.u.scan.key, W.sorted, G);
  double ih[5];
  inv_h_powers(I, ih);
  const int kt = A.kt;
  RowSums<Q, NF> S;
  S.zero();
  for (int c0 = 0; c0 < n; c0 += Q) {
    const int nc = min(Q, n - c0);
    wave_sync();  // the previous chunk's (or the sort's) reads come first
    if (G.gl < nc) {
      double a[Q], vw, df[NF];
      lsa_terms<Q, CLS>(A, I, me, W.sorted[c0 + G.gl], kt, ih, a, vw, df);
      double *row = W.u.rows[G.gl];
#pragma unroll
      for (int r = 0; r < Q; ++r) row[r] = a[r];
      row[RL::VW] = vw;
#pragma unroll
      for (int f = 0; f < NF; ++f) row[RL::DF + f] = df[f];
    }
    wave_sync();
    S.add_chunk(W, nc, G.gl);
  }
  wave_sync();
  double d[NF][5];
  if (n >= Q) {  // LSSPH_A_mod.f90:208 (num_neighbors >= q)
    row_dgesv<Q, NF>(W, S, G);
    double X[NF][5];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int k = 0; k < 5; ++k) X[f][k] = W.u.lu.r[k][Q + f];
    scale_derivs<NF>(X, I, d);
  } else {
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int k = 0; k < 5; ++k) d[f][k] = 0.0;
  }
  if (G.gl == 0) finish<CLS, true>(A, I, A.items[s], s, me, d);
}

// Particle shifting with q = 20 on row groups.
template <int Q, bool T>
__global__ void __launch_bounds__(64) k_pst_rows(PstArgs A) {
  constexpr int NF = T ? 4 : 3;  // u, v, rho (, tem)
  using RL = RowLds<Q, NF, 2>;
  constexpr int NG = RowGroup<Q>::G;
  __shared__ RL lds_all[NG];
  const RowGroup<Q> G(threadIdx.x);
  if (G.g >= NG) return;
  const int s = blockIdx.x * NG + G.g;
  if (s >= A.n) return;
  RL &W = lds_all[G.g];
  const int32_t inf = A.sinfo[s];
  const int ii = info_inst(inf);
  const InstDev &I = A.inst[ii];
  if (info_ptype(inf) != 0 || I.retired) {
    if (G.gl == 0) A.mout[s] = A.mir[s];  // walls move on with the records
    return;
  }
  const Rec me = A.mir[s];
  const int n = group_neighbors(A.mir, A.cstart, A.sinfo, A.items, A.orig_of, A.mv.counters, I, s, me,
                                W.u.scan.list, W.u.scan.key, W.sorted, G);
  double ih[5];
  inv_h_powers(I, ih);
  const int kt = A.kt;
  RowSums<Q, NF> S;
  S.zero();
  double sx = 0.0, sy = 0.0;
  for (int c0 = 0; c0 < n; c0 += Q) {
    const int nc = min(Q, n - c0);
    wave_sync();
    if (G.gl < nc) {
      double a[Q], vw, df[NF], shx, shy;
      pst_terms<Q, NF>(A, I, me, W.sorted[c0 + G.gl], kt, ih, a, vw, df, shx, shy);
      double *row = W.u.rows[G.gl];
#pragma unroll
      for (int r = 0; r < Q; ++r) row[r] = a[r];
      row[RL::VW] = vw;
#pragma unroll
      for (int f = 0; f < NF; ++f) row[RL::DF + f] = df[f];
      row[RL::XT] = shx;
      row[RL::XT + 1] = shy;
    }
    wave_sync();
    for (int k = 0; k < nc; ++k) {  // the shift sums in list order (:145-146)
      sx = sx + W.u.rows[k][RL::XT];
      sy = sy + W.u.rows[k][RL::XT + 1];
    }
    S.add_chunk(W, nc, G.gl);
  }
  wave_sync();
  // shift vector and its clamp (:217-225)
  const double umax = __longlong_as_double((long long)A.umax[ii]);
  const double coe = I.pst_c * umax * I.dt * I.h;
  double dxs = -coe * sx, dys = -coe * sy;
  const double rs = sqrt(dxs * dxs + dys * dys);
  if (rs > 0.2 * I.h) {
    dxs = (0.2 * I.h) * dxs / rs;
    dys = (0.2 * I.h) * dys / rs;
  }
  const bool solved = n >= Q;  // :240-247
  if (solved) row_dgesv<Q, NF>(W, S, G);
  if (G.gl != 0) return;
  double a[Q];
  basis_hq<Q>(dxs, dys, ih, a);
  double acc[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    double t = 0.0;
    if (solved) {
#pragma unroll
      for (int k = 0; k < Q; ++k) t += a[k] * W.u.lu.r[k][Q + f];
    }
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
// host launchers: groups of 64 / ppw lanes (ppw particles per wavefront)
// ---------------------------------------------------------------------------
template <int Q, int GS>
static void lsa_wave_g(hipStream_t s, const LsaArgs &a, int cls) {
  constexpr int per_block = kWavesPB * (64 / GS);
  const dim3 g((a.n + per_block - 1) / per_block), b(64 * kWavesPB);
  if (cls == 0)
    hipLaunchKernelGGL((k_lsa_wave<Q, 0, GS>), g, b, 0, s, a);
  else if (cls == 1)
    hipLaunchKernelGGL((k_lsa_wave<Q, 1, GS>), g, b, 0, s, a);
  else
    hipLaunchKernelGGL((k_lsa_wave<Q, 2, GS>), g, b, 0, s, a);
}

template <int Q>
static void lsa_rows(hipStream_t s, const LsaArgs &a, int cls) {
  constexpr int per_block = RowGroup<Q>::G;
  const dim3 g((a.n + per_block - 1) / per_block), b(64);
  if (cls == 0)
    hipLaunchKernelGGL((k_lsa_rows<Q, 0>), g, b, 0, s, a);
  else if (cls == 1)
    hipLaunchKernelGGL((k_lsa_rows<Q, 1>), g, b, 0, s, a);
  else
    hipLaunchKernelGGL((k_lsa_rows<Q, 2>), g, b, 0, s, a);
}

void launch_lsa_wave(hipStream_t s, Device &d, const LsaArgs &a, int, int cls) {
  if (a.n == 0) return;
  if (wave_particles(d) == 3)
    lsa_rows<20>(s, a, cls);
  else if (wave_particles(d) == 2)
    lsa_wave_g<20, 32>(s, a, cls);
  else
    lsa_wave_g<20, 64>(s, a, cls);
}

template <int Q, int GS>
static void pst_wave_g(hipStream_t s, const PstArgs &a, bool tem) {
  constexpr int per_block = kWavesPB * (64 / GS);
  const dim3 g((a.n + per_block - 1) / per_block), b(64 * kWavesPB);
  if (tem)
    hipLaunchKernelGGL((k_pst_wave<Q, true, GS>), g, b, 0, s, a);
  else
    hipLaunchKernelGGL((k_pst_wave<Q, false, GS>), g, b, 0, s, a);
}

template <int Q>
static void pst_rows(hipStream_t s, const PstArgs &a, bool tem) {
  constexpr int per_block = RowGroup<Q>::G;
  const dim3 g((a.n + per_block - 1) / per_block), b(64);
  if (tem)
    hipLaunchKernelGGL((k_pst_rows<Q, true>), g, b, 0, s, a);
  else
    hipLaunchKernelGGL((k_pst_rows<Q, false>), g, b, 0, s, a);
}

void launch_pst_wave(hipStream_t s, Device &d, const PstArgs &a, int) {
  if (a.n == 0) return;
  const bool tem = d.opt.has_tem != 0;
  if (wave_particles(d) == 3)
    pst_rows<20>(s, a, tem);
  else if (wave_particles(d) == 2)
    pst_wave_g<20, 32>(s, a, tem);
  else
    pst_wave_g<20, 64>(s, a, tem);
}

}  // namespace lssph
