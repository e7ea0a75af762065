// boundary.hip -- ghost-wall particles for gfx950 (ghost_mod.f90:43-78).
//
// A group of lanes per virtual marker (the mirror of a wall particle across
// the wall): interpolate the fluid values at the marker from its fluid
// neighbors in the cell-sorted mirror (LS type B with q = 3 or 6, or CSPH),
// then extrapolate them to the paired wall particle per wall type and
// boundary condition.  The wall values are written to the wall particle's
// slot of the current records (read by the next pair kernel).
//
// Several lanes per marker shorten each lane's serial chain of dependent
// loads and put more waves in flight when the markers are few; too many lose
// to the group sums and the repeated solves.  The launcher takes 2 lanes for
// the q = 3 fit of WALL_MODEL 2 (100 K markers of a 64-cavity ensemble: 1,
// 2, 4, 8 lanes take 69, 49, 51, 58 us per launch) and 4 for the others
// (16 K markers of 1024^2 Boussinesq convection, WALL_MODEL 3: 2, 4, 8
// lanes take 51, 31, 35 us).  The lane count sets the order in which a
// marker's partial sums are added; it follows from the wall model alone, a
// handle-wide option, so the wall values of an instance do not depend on
// the size of the ensemble it runs in.
#include "pair_common.hpp"

namespace lssph {

struct GhostArgs {
  Rec *__restrict__ mir;   // current records: fluid read, wall slots written
  float4 *__restrict__ nbf;  // their FP32 neighbor fields (mixed path; else null)
  const int32_t *__restrict__ sinfo;
  const int32_t *__restrict__ slot_of;
  const int32_t *__restrict__ cstart;
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

// Sum over the G lanes that share one virtual marker (G a power of two,
// groups aligned to G lanes).  Every lane of the group ends with the total.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int m = 1; m < G; m <<= 1) v += __shfl_xor(v, m, G);
  return v;
}

// Calls add(nb, x_j - x_m, y_j - y_m, r^2) for the fluid candidates t0,
// t0 + G, ... below t1 within h_eff of the marker (mx, my), in that order.
// With BATCH, four candidates at a time: their types and positions are
// loaded together, then the records of the accepted ones, so a lane waits on
// two load rounds per four candidates instead of up to two per candidate.
// That pays for the q = 6 fit (16 K markers of 1024^2 Boussinesq convection:
// 28.5 instead of 31 us per launch) and not for the lighter q = 3 fit (the
// 64-cavity ensemble's 100 K markers: 57 instead of 49 us, the batch's
// registers costing a wave per SIMD).  (The CSPH sums take the offsets'
// sign of interpolate_CSPH_mod.f90, x_m - x_j: r^2 is the same, and the
// Shepard sums use r only.)
template <int G, bool BATCH, typename F>
__device__ __forceinline__ void fluid_candidates(const GhostArgs &A, const InstDev &I, int t0, int t1, double mx,
                                                 double my, F &&add) {
  int t = t0;
  if constexpr (BATCH) {
    for (; t + 3 * G < t1; t += 4 * G) {
      int32_t inf[4];
      double2 xy[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        inf[j] = A.sinfo[t + j * G];
        xy[j] = load_xy(A.mir, t + j * G);
      }
      double xj[4], yj[4], r2[4];
      bool ok[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        xj[j] = xy[j].x - mx;
        yj[j] = xy[j].y - my;
        r2[j] = sq_dist(xj[j], yj[j]);
        ok[j] = info_ptype(inf[j]) == 0 && r2[j] <= I.h_eff_sq;
      }
      Rec nb[4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (ok[j]) nb[j] = A.mir[t + j * G];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (ok[j]) add(nb[j], xj[j], yj[j], r2[j]);
    }
  }
  for (; t < t1; t += G) {
    const int32_t inf = A.sinfo[t];
    const double2 xy = load_xy(A.mir, t);
    if (info_ptype(inf) != 0) continue;
    const double xj = xy.x - mx, yj = xy.y - my;
    const double r2 = sq_dist(xj, yj);
    if (r2 > I.h_eff_sq) continue;
    const Rec nb = A.mir[t];
    add(nb, xj, yj, r2);
  }
}

// interpolate_LSSPH_B_mod.f90:24-178: a = [1, x/h, y/h (, x^2/2h^2, xy/h^2,
// y^2/2h^2)], fluid neighbors only, value = constant term of the fit.  The
// G lanes of a marker take every G-th candidate of each contiguous cell run
// and add their partial sums at the end.
template <int Q, int G>
__device__ __forceinline__ void interp_lssph_b(const GhostArgs &A, const InstDev &I, int k, int lane, int kt,
                                               bool has_t, double (&val)[4]) {
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
  auto add = [&](const Rec &nb, double xj, double yj, double r2) {
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
  };
  visit_cells(I, A.vm_cell[k] - I.cell_off, A.cstart, [&](int t0, int t1, int, int) {
    fluid_candidates<G, Q == 6>(A, I, t0 + lane, t1, mx, my, add);
  });
  num = group_sum<G>(num);
#pragma unroll
  for (int i = 0; i < Q * (Q + 1) / 2; ++i) M[i] = group_sum<G>(M[i]);
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int i = 0; i < Q; ++i) b[f][i] = group_sum<G>(b[f][i]);
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

// interpolate_CSPH_mod.f90:24-130: Shepard average of the fluid neighbors,
// split over G lanes like interp_lssph_b.
template <int G>
__device__ __forceinline__ void interp_csph(const GhostArgs &A, const InstDev &I, int k, int lane, int kt,
                                            bool has_t, double (&val)[4]) {
  const double mx = A.vm_x[k], my = A.vm_y[k];
  double su = 0, sv = 0, sp = 0, st = 0, sw = 0;
  int num = 0;
  auto add = [&](const Rec &nb, double, double, double r2) {
    const double w = I.w_norm * kshape_w<double>(kt, sqrt(r2) * I.inv_h);
    ++num;
    const double vw = nb.vol * w;
    sw = sw + vw;
    su = su + vw * nb.u;
    sv = sv + vw * nb.v;
    sp = sp + vw * nb.pre;
    if (has_t) st = st + vw * nb.tem;
  };
  visit_cells(I, A.vm_cell[k] - I.cell_off, A.cstart, [&](int t0, int t1, int, int) {
    fluid_candidates<G, false>(A, I, t0 + lane, t1, mx, my, add);
  });
  num = group_sum<G>(num);
  sw = group_sum<G>(sw);
  su = group_sum<G>(su);
  sv = group_sum<G>(sv);
  sp = group_sum<G>(sp);
  st = group_sum<G>(st);
  if (num > 0 && sw > 1.0e-12) {
    val[0] = su / sw;
    val[1] = sv / sw;
    val[2] = sp / sw;
    val[3] = st / sw;
  } else {
    val[0] = val[1] = val[2] = val[3] = 0.0;
  }
}

// G lanes per virtual marker; the whole group leaves together when the
// marker index is past the end, so the group sums read active lanes only.
// WM is the wall model (1 CSPH, 2 LS q = 3, 3 LS q = 6): one instantiation
// each, so the q = 6 registers do not lower the occupancy of the others.
template <int G, int WM>
__global__ void __launch_bounds__(256) k_ghost(GhostArgs A) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = g / G, lane = g % G;
  if (k >= A.n_ext) return;
  const Options &o = A.opt;
  const InstDev &I = A.inst[A.vm_inst[k]];
  if (I.retired) return;  // a retired instance: the whole group leaves
  const bool has_t = o.has_tem != 0;
  double val[4];
  if constexpr (WM == 1)
    interp_csph<G>(A, I, k, lane, o.kernel_type, has_t, val);
  else if constexpr (WM == 3)
    interp_lssph_b<6, G>(A, I, k, lane, o.kernel_type, has_t, val);
  else
    interp_lssph_b<3, G>(A, I, k, lane, o.kernel_type, has_t, val);
  if (lane != 0) return;
  A.vm_u[k] = val[0];
  A.vm_v[k] = val[1];
  A.vm_pre[k] = val[2];
  if (has_t) A.vm_tem[k] = val[3];

  const int pw = A.vm_wall[k];
  const int sw = A.slot_of[pw];
  Rec w = A.mir[sw];
  const int t = A.vm_ptype[k];
  const double vu = val[0], vv = val[1], vp = val[2], vt = val[3];
  if (o.target_problem == 1) {
    // extrapolate_diff_test_mod.f90:38-104
    const double vx = A.vm_x[k], vy = A.vm_y[k];
    if (t == 1)
      w.u = (o.u_bc_bottom == 1) ? vu : 2.0 * (sin(kPi * (vx - I.wl_thick))) - vu;
    else if (t == 2)
      w.u = (o.u_bc_top == 1) ? vu : 2.0 * (-sin(kPi * (vx - I.wl_thick))) - vu;
    else if (t == 3)
      w.u = (o.u_bc_left == 1) ? vu + kPi * cos(kPi * (vy - I.wl_thick)) * (w.x - vx) : -vu;
    else if (t == 4)
      w.u = (o.u_bc_right == 1) ? vu + (-kPi) * cos(kPi * (vy - I.wl_thick)) * (w.x - vx) : -vu;
    else
      w.u = -vu;
  } else {
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
  store_rec(A.mir, A.nbf, sw, w);
}

// Virtual markers from the wall positions (init_virtual_markers_mod.f90:
// 61-137) and their fixed cells (init_cell_mod.f90:166-176).
__global__ void k_vm_init(int n_int, int n_ext, const Rec *__restrict__ rec, const int32_t *__restrict__ slot_of,
                          const int32_t *__restrict__ info,
                          const int32_t *__restrict__ phys_of, const InstDev *__restrict__ inst, double *vm_x,
                          double *vm_y, double *vm_nx, double *vm_ny, int32_t *vm_ptype, int32_t *vm_inst,
                          int32_t *vm_cell, int32_t *vm_wall, int32_t *escaped) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_ext) return;
  const int pw = phys_of[n_int + k];
  const int32_t inf = info[pw];
  const int t = info_ptype(inf), ii = info_inst(inf);
  const InstDev &I = inst[ii];
  const double wx = rec[slot_of[pw]].x, wy = rec[slot_of[pw]].y;
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
  int c = fixed_cell(I, x, y);
  if (t < 1 || t > 8 || c < 0) {
    atomicAdd(escaped, 1);
    c = 0;
  }
  vm_x[k] = x;
  vm_y[k] = y;
  vm_nx[k] = nx;
  vm_ny[k] = ny;
  vm_ptype[k] = t;
  vm_inst[k] = ii;
  vm_cell[k] = I.cell_off + c;
  vm_wall[k] = pw;
}

// ---------------------------------------------------------------------------
// Type-B LS fit of one field at probe points, the interpolation the
// reference's analysis uses for wall gradients (analysis/common/lssph_b.py:
// 21-144, called by analysis/benchmarks/boussinesq_conv.py:91-99 for the
// Nusselt number): every particle -- fluid and wall -- within h_eff of the
// probe, weights W(r) m / rho_j, basis a = [1, x/h, y/h (, x^2/2h^2, xy/h^2,
// y^2/2h^2 (, x^3/6h^3, x^2y/2h^3, xy^2/2h^3, y^3/6h^3))], offsets x = x_j -
// x_probe; out = (d_0, d_1 / h, d_2 / h), NaN when fewer than q particles
// are found or the moment matrix is singular.  FP64, one thread per probe.
// ---------------------------------------------------------------------------
struct ProbeArgs {
  const Rec *__restrict__ mir;
  const int32_t *__restrict__ cstart;
  const InstDev *__restrict__ inst;
  const double *__restrict__ px, *__restrict__ py;
  double *out;
  int n, inst_id, field, kt;
};

template <int Q>
__device__ __forceinline__ void probe_basis(double x, double y, double ih, double (&a)[Q]) {
  const double xh = x * ih, yh = y * ih;
  a[0] = 1.0;
  a[1] = xh;
  a[2] = yh;
  if constexpr (Q >= 6) {
    a[3] = 0.5 * xh * xh;
    a[4] = xh * yh;
    a[5] = 0.5 * yh * yh;
  }
  if constexpr (Q >= 10) {
    a[6] = (1.0 / 6.0) * xh * xh * xh;
    a[7] = 0.5 * xh * xh * yh;
    a[8] = 0.5 * xh * yh * yh;
    a[9] = (1.0 / 6.0) * yh * yh * yh;
  }
}

__device__ __forceinline__ double rec_field(const Rec &r, int f) {
  switch (f) {
    case 0: return r.u;
    case 1: return r.v;
    case 2: return r.rho;
    case 3: return r.pre;
    default: return r.tem;
  }
}

template <int Q>
__global__ void __launch_bounds__(64) k_probe_b(ProbeArgs A) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.n) return;
  const InstDev &I = A.inst[A.inst_id];
  const double mx = A.px[k], my = A.py[k];
  const double nan = __longlong_as_double(0x7ff8000000000000ll);
  const int c = fixed_cell(I, mx, my);
  if (c < 0) {
    A.out[3 * k] = A.out[3 * k + 1] = A.out[3 * k + 2] = nan;
    return;
  }
  double M[Q][Q], b[1][Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    b[0][i] = 0.0;
#pragma unroll
    for (int j = 0; j < Q; ++j) M[i][j] = 0.0;
  }
  int num = 0;
  const bool per = I.periodic;
  visit_cells(I, c, A.cstart, [&](int t0, int t1, int, int) {
    for (int t = t0; t < t1; ++t) {
      const Rec nb = A.mir[t];
      const double sx = per ? image_shift(nb.x, mx, I.len_x) : 0.0;
      const double sy = per ? image_shift(nb.y, my, I.len_y) : 0.0;
      const double xj = (nb.x + sx) - mx, yj = (nb.y + sy) - my;
      const double r2 = sq_dist(xj, yj);
      if (r2 > I.h_eff_sq) continue;
      ++num;
      const double vw = nb.vol * (I.w_norm * kshape_w<double>(A.kt, sqrt(r2) * I.inv_h));
      double a[Q];
      probe_basis<Q>(xj, yj, I.inv_h, a);
      const double f = rec_field(nb, A.field);
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const double vwa = vw * a[i];
        b[0][i] += vwa * f;
#pragma unroll
        for (int j = 0; j <= i; ++j) M[i][j] += vwa * a[j];
      }
    }
  });
#pragma unroll
  for (int i = 0; i < Q; ++i)
#pragma unroll
    for (int j = i + 1; j < Q; ++j) M[i][j] = M[j][i];
  if (num < Q || lu_solve<Q, 1>(M, b) != 0) {
    A.out[3 * k] = A.out[3 * k + 1] = A.out[3 * k + 2] = nan;
    return;
  }
  A.out[3 * k] = b[0][0];
  A.out[3 * k + 1] = b[0][1] * I.inv_h;
  A.out[3 * k + 2] = b[0][2] * I.inv_h;
}

static inline int blocks(int n, int b) { return (n + b - 1) / b; }

template <int G>
static void ghost_launch(hipStream_t s, int wall_model, int n_ext, const GhostArgs &a) {
  const dim3 grid(blocks(n_ext, 256 / G)), block(256);
  if (wall_model == 1)
    hipLaunchKernelGGL((k_ghost<G, 1>), grid, block, 0, s, a);
  else if (wall_model == 3)
    hipLaunchKernelGGL((k_ghost<G, 3>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_ghost<G, 2>), grid, block, 0, s, a);
}

void launch_ghost(hipStream_t s, Device &d, int cur) {
  if (d.n_ext == 0) return;
  GhostArgs a;
  a.mir = d.mir[cur];
  a.nbf = d.nbf[cur];
  a.sinfo = d.sinfo;
  a.slot_of = d.slot_of;
  a.cstart = d.cell_start;
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
  if (d.opt.wall_model == 2)
    ghost_launch<2>(s, d.opt.wall_model, d.n_ext, a);
  else
    ghost_launch<4>(s, d.opt.wall_model, d.n_ext, a);
}

void launch_probe(hipStream_t s, Device &d, int cur, int inst, int field, int q, int n, const double *px,
                  const double *py, double *out) {
  ProbeArgs a{d.mir[cur], d.cell_start, d.inst, px, py, out, n, inst, field, d.opt.kernel_type};
  const dim3 g(blocks(n, 64)), b(64);
  if (q == 3)
    hipLaunchKernelGGL(k_probe_b<3>, g, b, 0, s, a);
  else if (q == 6)
    hipLaunchKernelGGL(k_probe_b<6>, g, b, 0, s, a);
  else
    hipLaunchKernelGGL(k_probe_b<10>, g, b, 0, s, a);
}

void launch_vm_init(hipStream_t s, Device &d, int cur) {
  if (d.n_ext == 0) return;
  hipLaunchKernelGGL(k_vm_init, dim3(blocks(d.n_ext, 256)), dim3(256), 0, s, d.n_int, d.n_ext, d.mir[cur],
                     d.slot_of, d.info, d.phys_of, d.inst, d.vm_x, d.vm_y, d.vm_nx, d.vm_ny, d.vm_ptype, d.vm_inst,
                     d.vm_cell, d.vm_wall, d.counters + kEscaped);
}

}  // namespace lssph
