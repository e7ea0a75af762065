// transfer.hip -- host<->device transfers between the upload order and the
// slot order of the records, the neighbor lists dumped for validation, and
// the steady-state copy of the diffusion test.
#include "pair_common.hpp"

namespace lssph {

// Neighbor sets N(i) = { j != i in the 3x3 stencil, r_ij^2 <= h_eff^2 }
// (LSSPH_A_mod.f90:102-124) in summation order, as upload indices.
__global__ void k_nbr_count(int n, const Rec *__restrict__ mir, const int32_t *__restrict__ sinfo,
                            const int32_t *__restrict__ cstart, const int32_t *__restrict__ items,
                            const InstDev *__restrict__ inst, const int32_t *__restrict__ orig_of,
                            int32_t *counts_orig) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || info_ptype(sinfo[s]) != 0) return;
  const InstDev &I = inst[info_inst(sinfo[s])];
  const double2 me = load_xy(mir, s);
  int cnt = 0;
  visit_cells(I, local_cell(I, me.x, me.y), cstart, [&](int t0, int t1, int, int) {
    for (int t = t0; t < t1; ++t) {
      if (t == s) continue;
      const double2 xy = load_xy(mir, t);
      const double sx = I.periodic ? image_shift(xy.x, me.x, I.len_x) : 0.0;
      const double sy = I.periodic ? image_shift(xy.y, me.y, I.len_y) : 0.0;
      if (sq_dist((xy.x + sx) - me.x, (xy.y + sy) - me.y) <= I.h_eff_sq) ++cnt;
    }
  });
  counts_orig[orig_of[items[s]]] = cnt;
}

// The neighbors of each fluid particle in the reference's summation order:
// by position of the neighbor's reference cell in the 3x3 stencil of the
// particle's reference cell (rows dy = -1..1, columns dx = -1..1), then by
// upload index (fluid before walls, ascending).  The set is found on the
// internal cell grid and sorted in place (validation only).
__global__ void k_nbr_fill(int n, const Rec *__restrict__ mir, const int32_t *__restrict__ sinfo,
                           const int32_t *__restrict__ cstart, const int32_t *__restrict__ items,
                           const InstDev *__restrict__ inst, const int32_t *__restrict__ orig_of,
                           const int64_t *__restrict__ offsets, int32_t *list, int64_t *sort_key) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || info_ptype(sinfo[s]) != 0) return;
  const InstDev &I = inst[info_inst(sinfo[s])];
  const double2 me = load_xy(mir, s);
  int rcx, rcy;
  ref_cell_xy(I, me.x, me.y, false, rcx, rcy);
  const int64_t w0 = offsets[orig_of[items[s]]];
  int64_t w = w0;
  visit_cells(I, local_cell(I, me.x, me.y), cstart, [&](int t0, int t1, int, int) {
    for (int t = t0; t < t1; ++t) {
      if (t == s) continue;
      const double2 xy = load_xy(mir, t);
      const double sx = I.periodic ? image_shift(xy.x, me.x, I.len_x) : 0.0;
      const double sy = I.periodic ? image_shift(xy.y, me.y, I.len_y) : 0.0;
      if (sq_dist((xy.x + sx) - me.x, (xy.y + sy) - me.y) <= I.h_eff_sq) {
        int jx, jy;
        ref_cell_xy(I, xy.x, xy.y, info_ptype(sinfo[t]) != 0, jx, jy);
        int dx = jx - rcx, dy = jy - rcy;
        if (I.periodic) {
          dx = dx > 1 ? dx - I.rnx : (dx < -1 ? dx + I.rnx : dx);
          dy = dy > 1 ? dy - I.rny : (dy < -1 ? dy + I.rny : dy);
        }
        const int o = orig_of[items[t]];
        list[w] = o;
        sort_key[w] = ((int64_t)((dy + 1) * 3 + (dx + 1)) << 32) | (int64_t)o;
        ++w;
      }
    }
  });
  for (int64_t i = w0 + 1; i < w; ++i) {
    const int64_t k = sort_key[i];
    const int32_t v = list[i];
    int64_t j = i - 1;
    while (j >= w0 && sort_key[j] > k) {
      sort_key[j + 1] = sort_key[j];
      list[j + 1] = list[j];
      --j;
    }
    sort_key[j + 1] = k;
    list[j + 1] = v;
  }
}

// check_steady_state_mod.f90:50-67: flag per instance; instances that are
// not steady store u in f_previous.
__global__ void k_steady_copy(int n, const Rec *__restrict__ rec, const int32_t *__restrict__ slot_of,
                              const int32_t *__restrict__ info,
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
  if (!(df < inst[ii].threshold)) f_prev[p] = rec[slot_of[p]].u;
}

__global__ void k_download(int n, const Rec *__restrict__ rec, const int32_t *__restrict__ slot_of,
                           const int32_t *__restrict__ info, const int32_t *__restrict__ phys_of,
                           const InstDev *__restrict__ inst, double *stage, int32_t *ptype_out,
                           int32_t *cell_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = phys_of[i];
  const Rec r = rec[slot_of[p]];
  stage[0 * (size_t)n + i] = r.x;
  stage[1 * (size_t)n + i] = r.y;
  stage[2 * (size_t)n + i] = r.u;
  stage[3 * (size_t)n + i] = r.v;
  stage[4 * (size_t)n + i] = r.rho;
  stage[5 * (size_t)n + i] = r.pre;
  stage[6 * (size_t)n + i] = r.tem;
  const int pt = info_ptype(info[p]);
  ptype_out[i] = pt;
  cell_out[i] = ref_cell_index(inst[info_inst(info[p])], r.x, r.y, pt != 0);
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

// Records in upload order (slot = physical = upload order until the first
// cell build).
__global__ void k_upload(int n, const double *__restrict__ stage, const int32_t *__restrict__ ptype,
                         const int32_t *__restrict__ inst_of, const InstDev *__restrict__ inst, Rec *rec,
                         float4 *nbf, int32_t *info, int has_tem) {
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
  store_rec(rec, nbf, i, r);
  info[i] = make_info(ptype[i], ii);
}

static inline int blocks(int n, int b) { return (n + b - 1) / b; }

void launch_neighbor_count(hipStream_t s, Device &d, int cur, int32_t *counts_orig) {
  hipLaunchKernelGGL(k_nbr_count, dim3(blocks(d.n_total, 256)), dim3(256), 0, s, d.n_total, d.mir[cur], d.sinfo,
                     d.cell_start, d.items, d.inst, d.orig_of, counts_orig);
}

void launch_neighbor_fill(hipStream_t s, Device &d, int cur, const int64_t *offsets_orig, int32_t *list,
                          int64_t *sort_key) {
  hipLaunchKernelGGL(k_nbr_fill, dim3(blocks(d.n_total, 256)), dim3(256), 0, s, d.n_total, d.mir[cur], d.sinfo,
                     d.cell_start, d.items, d.inst, d.orig_of, offsets_orig, list, sort_key);
}

void launch_steady(hipStream_t s, Device &d, int cur) {
  launch_dfmax(s, d, cur);
  const int n = d.n_total > d.n_inst ? d.n_total : d.n_inst;
  hipLaunchKernelGGL(k_steady_copy, dim3(blocks(n, 256)), dim3(256), 0, s, d.n_total, d.mir[cur], d.slot_of,
                     d.info, d.inst, d.dfmax, d.f_prev, d.steady, d.n_inst);
}

void launch_download(hipStream_t s, Device &d, int cur) {
  const int n = d.n_total;
  hipLaunchKernelGGL(k_download, dim3(blocks(n, 256)), dim3(256), 0, s, n, d.mir[cur], d.slot_of, d.info,
                     d.phys_of, d.inst, d.stage_f64, d.stage_i32, d.perm_tmp);
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
                     d.mir[0], d.nbf[0], d.info, has_tem);
}

}  // namespace lssph
