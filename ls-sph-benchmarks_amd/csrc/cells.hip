// cells.hip -- cell-list builds for gfx950.
//
// The within-cell order is always the reference's single-thread cell-list
// order (update_cell_mod.f90:57-92: fluid in ascending index, then the fixed
// wall list in ascending index), i.e. ascending upload index, so pair sums
// are accumulated in the reference's order.  Two builds produce it:
//   full build (upload, re-sort, explicit update_cell): key + count (atomics
//     on ~10 particles per cell) -> three-kernel exclusive scan -> scatter ->
//     per-cell insertion sort;
//   incremental build (after every RK stage and shift): the kernel that
//     moved the particles has written each slot's new cell key; each new
//     cell gathers its members from the previous build's stencil of cells
//     (particles move far less than a cell per stage), counts them,
//     and after the scan writes them in order: the members that stayed are
//     already sorted, the few arrivals are merged in.  No atomics, no sort.
#include <climits>

#include "pair_common.hpp"

namespace lssph {

__global__ void k_fill_i32(int32_t *__restrict__ p, int32_t v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_iota(int32_t *__restrict__ p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

__global__ void k_add_i32(int32_t *__restrict__ p, int32_t v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p += v;
}

__global__ void k_fill_u64(unsigned long long *__restrict__ p, unsigned long long v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Lanes of a wavefront holding equal keys in consecutive lanes form runs
// (particles are kept roughly in cell order, so runs average several
// particles); one lane per run does the atomic for the whole run.
struct Run {
  int head_lane;  // first lane of my run
  int length;     // lanes in my run (valid on the head lane)
  bool head;
};

__device__ __forceinline__ Run lane_runs(int key) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1);
  const bool head = (lane == 0) || (prev != key);
  const unsigned long long hm = __ballot(head);
  const unsigned long long above = (lane == 63) ? 0ull : (hm >> (lane + 1)) << (lane + 1);
  const int next = above ? __ffsll((long long)above) - 1 : 64;
  const unsigned long long upto = (lane == 63) ? hm : (hm & ((2ull << lane) - 1ull));
  Run r;
  r.head = head;
  r.length = next - lane;
  r.head_lane = 63 - __clzll((long long)upto);
  return r;
}

// Keys of every particle and per-cell counts, over the slots of the current
// order (positions from the records in that order).  Fluid keys follow the
// particle positions; wall keys are fixed at upload (fixed_key_mode = 1
// computes them).  Escaped particles (outside the grid, or at a non-finite
// position) are counted as an error and spread over their instance's cells
// by slot, so that memory stays in bounds and no cell collects them.
__global__ void __launch_bounds__(256) k_bin_count(int n, const Rec *__restrict__ rec,
                                                   const int32_t *__restrict__ items,
                                                   const int32_t *__restrict__ info,
                                                   const InstDev *__restrict__ inst, int32_t *__restrict__ key,
                                                   int32_t *__restrict__ count, int32_t *__restrict__ escaped,
                                                   int fixed_key_mode) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  int k = -1 - (threadIdx.x & 63);  // distinct sentinel per idle lane
  if (t < n) {
    const int p = items[t];
    const int32_t inf = info[p];
    const InstDev &I = inst[info_inst(inf)];
    if (info_ptype(inf) == 0 || fixed_key_mode) {
      const double2 xy = *reinterpret_cast<const double2 *>(&rec[t].x);
      int c = info_ptype(inf) == 0 ? fluid_cell(I, xy.x, xy.y) : fixed_cell(I, xy.x, xy.y);
      if (c < 0) {
        atomicAdd(escaped, 1);
        c = t % (I.nx * I.ny);
      }
      k = I.cell_off + c;
      key[p] = k;
    } else {
      k = key[p];
    }
  }
  const Run r = lane_runs(k);
  if (r.head && k >= 0) atomicAdd(&count[k], r.length);
}

// --- exclusive scan of the cell counts ------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int block_exclusive_scan(int v, int *sh, int &total) {
  // Hillis-Steele over 256 lanes in LDS
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = 1; o < kScanThreads; o <<= 1) {
    int a = (t >= o) ? sh[t - o] : 0;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  total = sh[kScanThreads - 1];
  int ex = sh[t] - v;
  __syncthreads();
  return ex;
}

// Also folds the largest count into *max_out (one atomic per tile).
__global__ void __launch_bounds__(kScanThreads)
k_scan_tiles(const int32_t *__restrict__ in, int32_t *__restrict__ out, int32_t *__restrict__ tile_sums, int n,
             int32_t *__restrict__ max_out) {
  __shared__ int sh[kScanThreads];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0, mx = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < n) ? in[base + i] : 0;
    s += v[i];
    mx = max(mx, v[i]);
  }
  int total;
  int ex = block_exclusive_scan(s, sh, total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = ex;
    ex += v[i];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o));
  if (max_out && (threadIdx.x & 63) == 0) {
    sh[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (max_out && threadIdx.x == 0) {
    int b = sh[0];
    for (int w = 1; w < kScanThreads / 64; ++w) b = max(b, sh[w]);
    atomicMax(max_out, b);
  }
}

// One block scans the tile sums (in place, exclusive) and writes the grand
// total to out[n] (cell_start[n_cells] = number of particles).
__global__ void __launch_bounds__(kScanThreads)
k_scan_sums(int32_t *__restrict__ tile_sums, int ntiles, int32_t *__restrict__ grand) {
  __shared__ int sh[kScanThreads];
  int carry = 0;
  for (int base = 0; base < ntiles; base += kScanThreads) {
    const int i = base + threadIdx.x;
    int v = (i < ntiles) ? tile_sums[i] : 0;
    int total;
    int ex = block_exclusive_scan(v, sh, total);
    if (i < ntiles) tile_sums[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) *grand = carry;
}

__global__ void k_scan_add(int32_t *__restrict__ out, const int32_t *__restrict__ tile_sums, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += tile_sums[i / kScanTile];
}

// Scatter: a run reserves its slots by decrementing the cell count once,
// which also leaves count[] zero for the next build.  The order inside a
// cell is fixed afterwards by k_cell_sort.
__global__ void __launch_bounds__(256) k_scatter(int n, const int32_t *__restrict__ key,
                                                 const int32_t *__restrict__ start, int32_t *__restrict__ count,
                                                 int32_t *__restrict__ items) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int k = (p < n) ? key[p] : -1 - lane;
  const Run r = lane_runs(k);
  int base = 0;
  if (r.head && k >= 0) base = atomicSub(&count[k], r.length) - r.length;
  base = __shfl(base, r.head_lane);
  if (k >= 0) items[start[k] + base + (lane - r.head_lane)] = p;
}

// Per-cell sort of the slice by original particle index (insertion sort on
// the cell's <= kSortLds entries staged in LDS; larger cells sort in place
// in global memory).
constexpr int kSortThreads = 256;
constexpr int kSortLds = 16;

__global__ void __launch_bounds__(kSortThreads)
k_cell_sort(int ncells, const int32_t *__restrict__ start, int32_t *__restrict__ items,
            const int32_t *__restrict__ orig_of) {
  __shared__ int32_t sh_item[kSortLds][kSortThreads];
  __shared__ int32_t sh_key[kSortLds][kSortThreads];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int tid = threadIdx.x;
  const int s = (c < ncells) ? start[c] : 0, e = (c < ncells) ? start[c + 1] : 0;
  const int m = e - s;
  if (m > 1 && m <= kSortLds) {
    for (int i = 0; i < m; ++i) {
      const int it = items[s + i];
      const int ko = orig_of[it];
      int j = i - 1;
      while (j >= 0 && sh_key[j][tid] > ko) {
        sh_key[j + 1][tid] = sh_key[j][tid];
        sh_item[j + 1][tid] = sh_item[j][tid];
        --j;
      }
      sh_key[j + 1][tid] = ko;
      sh_item[j + 1][tid] = it;
    }
    for (int i = 0; i < m; ++i) items[s + i] = sh_item[i][tid];
  } else if (m > kSortLds) {
    for (int i = s + 1; i < e; ++i) {
      const int it = items[i];
      const int ko = orig_of[it];
      int j = i - 1;
      while (j >= s && orig_of[items[j]] > ko) {
        items[j + 1] = items[j];
        --j;
      }
      items[j + 1] = it;
    }
  }
}

// --- incremental build -------------------------------------------------------
// Members of new cell c: slots t of the previous build's stencil around c
// whose new key is c, plus listed strays with key c.  Members that were not
// in c before (arrivals) are recorded, up to kArrivals per cell, for the
// fill; narr[c] holds their number.
__global__ void __launch_bounds__(256)
k_cell_count(int ncells, const int32_t *__restrict__ cell_inst, const InstDev *__restrict__ inst,
             const int32_t *__restrict__ cstart, const int32_t *__restrict__ kslot,
             const int32_t *__restrict__ strays, const int32_t *__restrict__ nstray, int32_t *__restrict__ count,
             int32_t *__restrict__ narr, int32_t *__restrict__ arr) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const InstDev &I = inst[cell_inst[c]];
  const int o0 = cstart[c], o1 = cstart[c + 1];
  int cnt = 0, na = 0;
  auto arrival = [&](int t) {
    if (na < kArrivals) arr[(size_t)kArrivals * c + na] = t;
    ++na;
  };
  visit_cells(I, c - I.cell_off, cstart, [&](int t0, int t1, int, int) {
    int t = t0;
    for (; t + 4 <= t1; t += 4) {
      const int k0 = kslot[t], k1 = kslot[t + 1], k2 = kslot[t + 2], k3 = kslot[t + 3];
      if (k0 == c) { ++cnt; if (t < o0 || t >= o1) arrival(t); }
      if (k1 == c) { ++cnt; if (t + 1 < o0 || t + 1 >= o1) arrival(t + 1); }
      if (k2 == c) { ++cnt; if (t + 2 < o0 || t + 2 >= o1) arrival(t + 2); }
      if (k3 == c) { ++cnt; if (t + 3 < o0 || t + 3 >= o1) arrival(t + 3); }
    }
    for (; t < t1; ++t)
      if (kslot[t] == c) { ++cnt; if (t < o0 || t >= o1) arrival(t); }
  });
  const int ns = min(*nstray, kStrayCap);
  for (int k = 0; k < ns; ++k)
    if (strays[2 * k + 1] == c) {
      ++cnt;
      arrival(strays[2 * k]);
    }
  count[c] = cnt;
  narr[c] = na;
}

constexpr int kFillThreads = 128;
constexpr int kFillBatch = 16;  // previous-build slots loaded together per batch

// Writes the members of new cell c into its slots in ascending upload index.
// Members that stayed in c keep their (sorted) order from the previous
// build, each shifted by the number of arrivals before it; the recorded
// arrivals are sorted and placed after the stayers before them.  A cell with
// more than kArrivals arrivals selects its members one by one instead
// (correct for any count, slow, rare).  The previous build's slots are read
// in batches of kFillBatch independent loads, so a cell costs a couple of
// memory round trips rather than one per member.  Arrivals get their new
// key; slots are recorded by the mirror pass.  Clears count[c] for the next
// full build's atomics.
__global__ void __launch_bounds__(kFillThreads)
k_cell_fill(int ncells, const int32_t *__restrict__ cell_inst, const InstDev *__restrict__ inst,
            const int32_t *__restrict__ cstart_old, const int32_t *__restrict__ cstart_new,
            const int32_t *__restrict__ kslot, const int32_t *__restrict__ sorig,
            const int32_t *__restrict__ strays, const int32_t *__restrict__ nstray,
            const int32_t *__restrict__ narr, const int32_t *__restrict__ arr,
            const int32_t *__restrict__ items_old, int32_t *__restrict__ items_new, int32_t *__restrict__ key,
            int32_t *__restrict__ count) {
  __shared__ int32_t arr_t[kArrivals][kFillThreads];
  __shared__ int32_t arr_o[kArrivals][kFillThreads];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int tid = threadIdx.x;
  if (c >= ncells) return;
  count[c] = 0;
  const int base = cstart_new[c], m = cstart_new[c + 1] - base;
  if (m == 0) return;
  const int o0 = cstart_old[c], o1 = cstart_old[c + 1];
  const int na = narr[c];
  if (na > kArrivals) {
    // many arrivals: select members in ascending upload index
    const InstDev &I = inst[cell_inst[c]];
    const int lc = c - I.cell_off;
    const int ns = min(*nstray, kStrayCap);
    int last = -1;
    for (int k = 0; k < m; ++k) {
      int best = INT_MAX, bt = -1;
      visit_cells(I, lc, cstart_old, [&](int t0, int t1, int, int) {
        for (int t = t0; t < t1; ++t)
          if (kslot[t] == c) {
            const int o = sorig[t];
            if (o > last && o < best) {
              best = o;
              bt = t;
            }
          }
      });
      for (int q = 0; q < ns; ++q)
        if (strays[2 * q + 1] == c) {
          const int t = strays[2 * q], o = sorig[t];
          if (o > last && o < best) {
            best = o;
            bt = t;
          }
        }
      if (bt < 0) break;
      const int p = items_old[bt];
      items_new[base + k] = p;
      key[p] = c;
      last = best;
    }
    return;
  }
  // arrivals sorted by upload index (insertion sort in LDS)
  for (int i = 0; i < na; ++i) {
    const int t = arr[(size_t)kArrivals * c + i], o = sorig[t];
    int j = i - 1;
    while (j >= 0 && arr_o[j][tid] > o) {
      arr_o[j + 1][tid] = arr_o[j][tid];
      arr_t[j + 1][tid] = arr_t[j][tid];
      --j;
    }
    arr_o[j + 1][tid] = o;
    arr_t[j + 1][tid] = t;
  }
  // stayers: rank = stayers before + arrivals with a smaller index
  int before[kArrivals];
#pragma unroll
  for (int j = 0; j < kArrivals; ++j) before[j] = 0;
  int k = 0;
  for (int b = o0; b < o1; b += kFillBatch) {
    int ks[kFillBatch], it[kFillBatch], so[kFillBatch];
#pragma unroll
    for (int u = 0; u < kFillBatch; ++u) ks[u] = (b + u < o1) ? kslot[b + u] : -1;
#pragma unroll
    for (int u = 0; u < kFillBatch; ++u) {
      const bool stay = ks[u] == c;
      it[u] = stay ? items_old[b + u] : 0;
      so[u] = (stay && na > 0) ? sorig[b + u] : 0;
    }
#pragma unroll
    for (int u = 0; u < kFillBatch; ++u) {
      if (ks[u] != c) continue;
      int r = k;
      if (na > 0) {
#pragma unroll
        for (int j = 0; j < kArrivals; ++j)
          if (j < na) {
            const bool lt = arr_o[j][tid] < so[u];
            r += lt;
            before[j] += !lt;
          }
      }
      items_new[base + r] = it[u];
      ++k;
    }
  }
  // arrival j: rank = j + stayers with a smaller index
#pragma unroll
  for (int j = 0; j < kArrivals; ++j)
    if (j < na) {
      const int p = items_old[arr_t[j][tid]];
      items_new[base + j + before[j]] = p;
      key[p] = c;
    }
}

// Records in the new slot order: dst[t] = src[previous slot of the particle
// at t] (and their FP32 neighbor fields when nbf_dst is non-null), with the
// info, cell key and upload index of each slot and the position the cell
// list is built at.  slot_of[p] turns from the previous slot into the new one
// (each thread reads and writes only its own p).  Also empties the stray
// list for the next moves.
__global__ void k_mirror(int n, const int32_t *__restrict__ items, int32_t *__restrict__ slot_of,
                         const int32_t *__restrict__ info, const int32_t *__restrict__ key,
                         const int32_t *__restrict__ orig_of, const Rec *__restrict__ src, Rec *__restrict__ dst,
                         float4 *__restrict__ nbf_dst, int32_t *__restrict__ sinfo, int32_t *__restrict__ kslot,
                         int32_t *__restrict__ sorig, double *__restrict__ xbuild, int32_t *__restrict__ nstray) {
  const int t = xcd_block() * blockDim.x + threadIdx.x;
  if (t == 0) *nstray = 0;
  if (t >= n) return;
  const int p = items[t];
  const Rec r = src[slot_of[p]];
  store_rec(dst, nbf_dst, t, r);
  sinfo[t] = info[p];
  kslot[t] = key[p];
  sorig[t] = orig_of[p];
  slot_of[p] = t;
  xbuild[2 * t] = r.x;
  xbuild[2 * t + 1] = r.y;
}

// --- permutations (physical re-sort) --------------------------------------
__global__ void k_gather_i32(int n, const int32_t *__restrict__ perm, const int32_t *__restrict__ src,
                             int32_t *__restrict__ dst) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[perm[t]];
}

__global__ void k_gather_f64(int n, const int32_t *__restrict__ perm, const double *__restrict__ src,
                             double *__restrict__ dst) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[perm[t]];
}

__global__ void k_inverse_perm(int n, const int32_t *__restrict__ orig_of, int32_t *__restrict__ phys_of) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) phys_of[orig_of[t]] = t;
}

// --- host launchers ---------------------------------------------------------
static inline int blocks(int n, int b) { return (n + b - 1) / b; }

void launch_add_i32(hipStream_t s, int32_t *p, int32_t v) {
  hipLaunchKernelGGL(k_add_i32, dim3(1), dim3(64), 0, s, p, v);
}

void launch_fill_i32(hipStream_t s, int32_t *p, int32_t v, int n) {
  if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3(blocks(n, 256)), dim3(256), 0, s, p, v, n);
}

void launch_iota(hipStream_t s, int32_t *p, int n) {
  if (n > 0) hipLaunchKernelGGL(k_iota, dim3(blocks(n, 256)), dim3(256), 0, s, p, n);
}

void launch_fill_u64(hipStream_t s, unsigned long long *p, unsigned long long v, int n) {
  if (n > 0) hipLaunchKernelGGL(k_fill_u64, dim3(blocks(n, 256)), dim3(256), 0, s, p, v, n);
}

int scan_tile_count(int n) { return blocks(n, kScanTile); }

void launch_exclusive_scan(hipStream_t s, const int32_t *in, int32_t *out, int32_t *tile_sums, int n,
                           int32_t *max_out) {
  const int nt = scan_tile_count(n);
  hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kScanThreads), 0, s, in, out, tile_sums, n, max_out);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, tile_sums, nt, out + n);
  hipLaunchKernelGGL(k_scan_add, dim3(blocks(n, 256)), dim3(256), 0, s, out, tile_sums, n);
}

static void launch_mirror(hipStream_t s, Device &d, int src, int dst) {
  hipLaunchKernelGGL(k_mirror, dim3(xcd_grid(d.n_total, 256)), dim3(256), 0, s, d.n_total, d.items, d.slot_of,
                     d.info, d.key, d.orig_of, d.mir[src], d.mir[dst], d.nbf[dst], d.sinfo, d.kslot, d.sorig,
                     d.xbuild, d.nstray);
}

void launch_bin(hipStream_t s, Device &d, int src, int dst, int fixed_key_mode) {
  const int n = d.n_total;
  hipLaunchKernelGGL(k_bin_count, dim3(blocks(n, 256)), dim3(256), 0, s, n, d.mir[src], d.items, d.info, d.inst,
                     d.key, d.cell_count, d.counters + kEscaped, fixed_key_mode);
  launch_exclusive_scan(s, d.cell_count, d.cell_start, d.scan_tmp, d.n_cells, d.counters + kMaxCell);
  hipLaunchKernelGGL(k_scatter, dim3(blocks(n, 256)), dim3(256), 0, s, n, d.key, d.cell_start,
                     d.cell_count, d.items);
  hipLaunchKernelGGL(k_cell_sort, dim3(blocks(d.n_cells, kSortThreads)), dim3(kSortThreads), 0, s, d.n_cells,
                     d.cell_start, d.items, d.orig_of);
  launch_mirror(s, d, src, dst);
}

void select_build(Device &d, int ib) {
  d.ib = ib;
  d.items = d.items_buf[ib];
  d.cell_start = d.cstart_buf[ib];
}

// Incremental build after moves (the last mover pass has written kslot and
// the stray list); the next build's buffers become current.
void launch_rebuild(hipStream_t s, Device &d, int src, int dst) {
  const int nxt = d.ib ^ 1;
  hipLaunchKernelGGL(k_cell_count, dim3(blocks(d.n_cells, 256)), dim3(256), 0, s, d.n_cells, d.cell_inst, d.inst,
                     d.cell_start, d.kslot, d.strays, d.nstray, d.cell_count, d.narr, d.arr);
  launch_exclusive_scan(s, d.cell_count, d.cstart_buf[nxt], d.scan_tmp, d.n_cells, d.counters + kMaxCell);
  hipLaunchKernelGGL(k_cell_fill, dim3(blocks(d.n_cells, kFillThreads)), dim3(kFillThreads), 0, s, d.n_cells,
                     d.cell_inst, d.inst, d.cell_start, d.cstart_buf[nxt], d.kslot, d.sorig, d.strays, d.nstray,
                     d.narr, d.arr, d.items, d.items_buf[nxt], d.key, d.cell_count);
  select_build(d, nxt);
  launch_mirror(s, d, src, dst);
}

void launch_gather_i32(hipStream_t s, int n, const int32_t *perm, const int32_t *src, int32_t *dst) {
  hipLaunchKernelGGL(k_gather_i32, dim3(blocks(n, 256)), dim3(256), 0, s, n, perm, src, dst);
}

void launch_gather_f64(hipStream_t s, int n, const int32_t *perm, const double *src, double *dst) {
  hipLaunchKernelGGL(k_gather_f64, dim3(blocks(n, 256)), dim3(256), 0, s, n, perm, src, dst);
}

void launch_inverse_perm(hipStream_t s, int n, const int32_t *orig_of, int32_t *phys_of) {
  hipLaunchKernelGGL(k_inverse_perm, dim3(blocks(n, 256)), dim3(256), 0, s, n, orig_of, phys_of);
}

}  // namespace lssph
