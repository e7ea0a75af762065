// cells.hip -- cell-list build for gfx950: a counting sort of particles by
// cell, made deterministic by sorting each cell's slice by original index.
//
// The within-cell order is the reference's single-thread cell-list order
// (update_cell_mod.f90:57-92: fluid in ascending index, then the fixed wall
// list in ascending index), so pair sums are accumulated in the reference's
// order.  Passes: key + count (atomics on ~10 particles per cell) -> three
// kernel exclusive scan -> scatter -> per-cell insertion sort.
#include "pair_common.hpp"

namespace lssph {

__global__ void k_fill_i32(int32_t *__restrict__ p, int32_t v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
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

// Keys of every particle and per-cell counts.  Fluid keys follow the
// particle positions; wall keys are fixed at init (fixed_key_mode = 1
// computes them).  Escaped particles are counted and clamped into the grid
// so that memory stays in bounds.
__global__ void __launch_bounds__(256) k_bin_count(int n, const Rec *__restrict__ rec,
                                                   const int32_t *__restrict__ info,
                                                   const InstDev *__restrict__ inst, int32_t *__restrict__ key,
                                                   int32_t *__restrict__ count, int32_t *__restrict__ escaped,
                                                   int fixed_key_mode) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  int k = -1 - (threadIdx.x & 63);  // distinct sentinel per idle lane
  if (p < n) {
    const int32_t inf = info[p];
    const InstDev &I = inst[info_inst(inf)];
    if (info_ptype(inf) == 0 || fixed_key_mode) {
      const double2 xy = *reinterpret_cast<const double2 *>(&rec[p].x);
      int c = info_ptype(inf) == 0 ? fluid_cell(I, xy.x, xy.y) : fixed_cell(I, xy.x, xy.y);
      if (c < 0) {
        atomicAdd(escaped, 1);
        c = 0;
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

__global__ void __launch_bounds__(kScanThreads)
k_scan_tiles(const int32_t *__restrict__ in, int32_t *__restrict__ out, int32_t *__restrict__ tile_sums, int n) {
  __shared__ int sh[kScanThreads];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < n) ? in[base + i] : 0;
    s += v[i];
  }
  int total;
  int ex = block_exclusive_scan(s, sh, total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = ex;
    ex += v[i];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
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
// in global memory).  Also records the largest cell count.
constexpr int kSortThreads = 256;
constexpr int kSortLds = 16;

__global__ void __launch_bounds__(kSortThreads)
k_cell_sort(int ncells, const int32_t *__restrict__ start, int32_t *__restrict__ items,
            const int32_t *__restrict__ orig_of, int32_t *__restrict__ max_count) {
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
  int mm = m;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mm = max(mm, __shfl_xor(mm, o));
  if ((threadIdx.x & 63) == 0) atomicMax(max_count, mm);
}

// Cell-sorted mirror of the records: mirror[t] = rec[items[t]], the info of
// each slot, and the slot of every physical particle.
__global__ void k_mirror(int n, const int32_t *__restrict__ items, const Rec *__restrict__ rec,
                         const int32_t *__restrict__ info, Rec *__restrict__ mirror, int32_t *__restrict__ sinfo,
                         int32_t *__restrict__ slot_of) {
  const int t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int p = items[t];
  mirror[t] = rec[p];
  sinfo[t] = info[p];
  slot_of[p] = t;
}

// --- permutations (physical re-sort) --------------------------------------
__global__ void k_gather_rec(int n, const int32_t *__restrict__ perm, const Rec *__restrict__ src,
                             Rec *__restrict__ dst) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[perm[t]];
}

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

void launch_fill_i32(hipStream_t s, int32_t *p, int32_t v, int n) {
  if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3(blocks(n, 256)), dim3(256), 0, s, p, v, n);
}

void launch_fill_u64(hipStream_t s, unsigned long long *p, unsigned long long v, int n) {
  if (n > 0) hipLaunchKernelGGL(k_fill_u64, dim3(blocks(n, 256)), dim3(256), 0, s, p, v, n);
}

int scan_tile_count(int n) { return blocks(n, kScanTile); }

void launch_exclusive_scan(hipStream_t s, const int32_t *in, int32_t *out, int32_t *tile_sums, int n) {
  const int nt = scan_tile_count(n);
  hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kScanThreads), 0, s, in, out, tile_sums, n);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, tile_sums, nt, out + n);
  hipLaunchKernelGGL(k_scan_add, dim3(blocks(n, 256)), dim3(256), 0, s, out, tile_sums, n);
}

void launch_bin(hipStream_t s, Device &d, const Rec *rec, int fixed_key_mode) {
  const int n = d.n_total;
  hipLaunchKernelGGL(k_bin_count, dim3(blocks(n, 256)), dim3(256), 0, s, n, rec, d.info, d.inst, d.key,
                     d.cell_count, d.counters + kEscaped, fixed_key_mode);
  launch_exclusive_scan(s, d.cell_count, d.cell_start, d.scan_tmp, d.n_cells);
  hipLaunchKernelGGL(k_scatter, dim3(blocks(n, 256)), dim3(256), 0, s, n, d.key, d.cell_start,
                     d.cell_count, d.items);
  hipLaunchKernelGGL(k_cell_sort, dim3(blocks(d.n_cells, kSortThreads)), dim3(kSortThreads), 0, s, d.n_cells,
                     d.cell_start, d.items, d.orig_of, d.counters + kMaxCell);
  launch_mirror(s, d, rec);
}

void launch_mirror(hipStream_t s, Device &d, const Rec *rec) {
  hipLaunchKernelGGL(k_mirror, dim3(xcd_grid(d.n_total, 256)), dim3(256), 0, s, d.n_total, d.items, rec, d.info,
                     d.mirror, d.sinfo, d.slot_of);
}

void launch_gather_rec(hipStream_t s, int n, const int32_t *perm, const Rec *src, Rec *dst) {
  hipLaunchKernelGGL(k_gather_rec, dim3(blocks(n, 256)), dim3(256), 0, s, n, perm, src, dst);
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
