// internal.hpp -- device-side state of a handle and the kernel launchers
// shared by the translation units of liblssph_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/lssph.h"
#include "device_common.hpp"

namespace lssph {

// Indices into Device::counters (int32) and Device::stats64 (uint64).
enum {
  kEscaped = 0,
  kMaxCell = 1,
  kMaxNbr = 2,
  kFbCount = 3,
  kStrayOverflow = 4,
  kSkinExceeded = 5,
  kWaveListOverflow = 6,  // a particle of the q = 20 kernels with too many neighbors
  kFbDone = 7,            // blocks of the fallback kernel done (the last one clears the queue)
  kStepNo = 8,            // isolation mode: 1-based number of the current time step
  kNumCounters = 9
};

// Neighbor-list entries of the per-lane scan (and of the q = 20 kernels) are
// 32 bits: the slot in the low kSlotBits bits, the periodic image code above
// it.  lssph_create therefore refuses handles of more than kMaxParticles
// particles (slots 0 .. kSlotMask).
constexpr int kSlotBits = 27;
constexpr int kSlotMask = (1 << kSlotBits) - 1;
constexpr long long kMaxParticles = kSlotMask;

// Incremental cell rebuild: particles whose new cell is outside the
// stencil of their old cell are listed (old slot, new key) for the rebuild;
// more than kStrayCap of them in one stage is reported as an error.
constexpr int kStrayCap = 4096;
constexpr int kArrivals = 8;  // arrivals per new cell merged directly
// Neighbor-list capacity per particle of the lane-group kernels (SPH_MODEL
// 6).
constexpr int kWaveList = 128;
enum { kFallbacks = 0, kOverflows = 1, kNumStats64 = 2 };

// Raw device pointers and sizes; passed to the launchers.
struct Device {
  int n_total = 0, n_int = 0, n_ext = 0, n_inst = 0, n_cells = 0;
  Options opt{};
  InstDev *inst = nullptr;          // [n_inst]
  Rec *mir[2] = {nullptr, nullptr};  // records in slot order, double-buffered across passes
  float4 *nbf[2] = {nullptr, nullptr};  // mixed path only: nb_fields of mir[b], same slots
  int32_t *info = nullptr;          // [n_total] ptype | inst << 8
  int32_t *key = nullptr;           // [n_total] global cell id
  int32_t *orig_of = nullptr;       // [n_total] physical -> upload index
  int32_t *phys_of = nullptr;       // [n_total] upload index -> physical
  int32_t *cell_count = nullptr;    // [n_cells]
  int32_t *cell_start = nullptr;    // [n_cells + 1] current build (points into cstart_buf)
  int32_t *items = nullptr;         // [n_total] slot -> physical id, current build (items_buf)
  int32_t *cstart_buf[2] = {};      // cell starts of the current and the next build
  int32_t *items_buf[2] = {};
  int ib = 0;                       // index of the current build's buffers
  int32_t *cell_inst = nullptr;     // [n_cells] instance of each cell
  int32_t *kslot = nullptr;         // [n_total] cell key per slot (movers write the new key)
  int32_t *strays = nullptr;        // [2 * kStrayCap] (old slot, new key)
  int32_t *nstray = nullptr;        // [1]
  int32_t *sorig = nullptr;         // [n_total] upload index per slot
  int32_t *narr = nullptr;          // [n_cells] arrivals per new cell
  int32_t *arr = nullptr;           // [kArrivals * n_cells] their previous slots
  double *xbuild = nullptr;         // [2 * n_total] position per slot at the last cell build
  int32_t *sinfo = nullptr;         // [n_total] info in slot order
  int32_t *slot_of = nullptr;       // [n_total] physical id -> slot
  int32_t *fb_list = nullptr;       // [n_total] slots queued for the FP64 fallback
  int32_t *scan_tmp = nullptr;      // scan tile sums
  int32_t *perm_tmp = nullptr;      // [n_total]
  // RK work arrays (physical order, fluid entries used)
  double *old_[6] = {};  // x, y, u, v, rho, tem
  double *new_[6] = {};
  double *rhs[4] = {};   // u, v, rho, tem (lssph_calculate_rhs only)
  double *f_prev = nullptr;  // diffusion steady check
  double *spd = nullptr;     // per slot: |u| after the last RK stage (0 for walls), for U_max
  // virtual markers (marker k pairs with upload index n_int + k)
  double *vm_x = nullptr, *vm_y = nullptr, *vm_nx = nullptr, *vm_ny = nullptr;
  double *vm_u = nullptr, *vm_v = nullptr, *vm_pre = nullptr, *vm_tem = nullptr;
  int32_t *vm_ptype = nullptr, *vm_inst = nullptr, *vm_cell = nullptr, *vm_wall = nullptr;
  // reductions and counters
  unsigned long long *umax = nullptr;    // [n_inst] bits of max |u|
  unsigned long long *dfmax = nullptr;   // [n_inst] bits of max |f_prev - u|
  int32_t *steady = nullptr;             // [n_inst]
  int32_t *counters = nullptr;           // [kNumCounters]
  unsigned long long *stats64 = nullptr;  // [kNumStats64]
  // isolation mode (lssph_set_isolation): retired[k] = step during which
  // instance k was retired, 0 while it runs, retired[n_inst + k] = reason;
  // null when the mode is off
  int32_t *retired = nullptr;             // [2 n_inst]
  int32_t *retired_buf = nullptr;         // the allocation behind retired
  // host-side staging for downloads (device scratch, upload order)
  double *stage_f64 = nullptr;  // [8 * n_total]
  int32_t *stage_i32 = nullptr;  // [n_total]
  // tuning
  int block = 128;
  int staged_scan = 1;  // lssph_set_scan
  int list_len = 39;     // LDS list entries per thread, the particle itself included
  int wave_ppw = 0;      // particles per wavefront of the SPH_MODEL 6 kernels (0: default)
  int red_blocks = 256;  // blocks per instance of the max reductions
};

// The mixed-precision path runs (and the FP32 neighbor fields nbf are kept)
// for SPH_MODEL 3 handles created with precision LSSPH_MIXED.
inline bool mixed_path(const Device &d) { return d.opt.precision == LSSPH_MIXED && d.opt.sph_model == 3; }

// Particles per wavefront of the SPH_MODEL 6 kernels: 1 or 2 (lane groups
// of 64 or 32), 3 (row groups of 20 lanes); the default is the fastest
// setting measured at 1024^2.
inline int wave_particles(const Device &d) { return d.wave_ppw ? d.wave_ppw : 3; }

// cells.hip
void launch_fill_i32(hipStream_t s, int32_t *p, int32_t v, int n);
void launch_fill_u64(hipStream_t s, unsigned long long *p, unsigned long long v, int n);
void launch_add_i32(hipStream_t s, int32_t *p, int32_t v);  // *p += v (one thread)
int scan_tile_count(int n);
void launch_exclusive_scan(hipStream_t s, const int32_t *in, int32_t *out, int32_t *tile_sums, int n,
                           int32_t *max_out);
// Cell builds read the records of buffer src (in the previous build's slot
// order) and write them, in the new order, to buffer dst.
void launch_bin(hipStream_t s, Device &d, int src, int dst, int fixed_key_mode);
void launch_rebuild(hipStream_t s, Device &d, int src, int dst);
void select_build(Device &d, int ib);
void launch_iota(hipStream_t s, int32_t *p, int n);
void launch_gather_i32(hipStream_t s, int n, const int32_t *perm, const int32_t *src, int32_t *dst);
void launch_gather_f64(hipStream_t s, int n, const int32_t *perm, const double *src, double *dst);
void launch_inverse_perm(hipStream_t s, int n, const int32_t *orig_of, int32_t *phys_of);

// operators.hip
// Stage codes of the fused LS type-A kernel.
enum StageMode { kRhsOnly = 0, kRk2S1, kRk2S2, kRk4S1, kRk4S2, kRk4S3, kRk4S4 };
// Mover passes read records buffer cur and write buffer cur ^ 1; guard
// checks each move against the skin, keys prepares the next cell build.
// write_spd: the last RK stage also stores each particle's speed for U_max.
void launch_lsa(hipStream_t s, Device &d, int cur, StageMode mode, int guard, int keys, bool write_spd = false);
// U_max from the speeds the last RK stage stored (valid while slots are unchanged).
void launch_umax_spd(hipStream_t s, Device &d);
void launch_umax(hipStream_t s, Device &d, int cur);
// The neighbor lists of the pair kernels as their walk visits them (upload
// indices, cap entries per fluid particle) and the scan path of each.
void launch_pair_lists(hipStream_t s, Device &d, int cur, int32_t *counts, int32_t *list, int32_t *path, int cap);
void launch_dfmax(hipStream_t s, Device &d, int cur);
void launch_pst(hipStream_t s, Device &d, int cur, int guard, int keys);
void launch_ghost(hipStream_t s, Device &d, int cur);
void launch_vm_init(hipStream_t s, Device &d, int cur);
// Type-B LS fit of field (0 u, 1 v, 2 rho, 3 pre, 4 tem) at n probe points of
// one instance, q = 3, 6 or 10: out[3k..3k+2] = value, d/dx, d/dy.
void launch_probe(hipStream_t s, Device &d, int cur, int inst, int field, int q, int n, const double *px,
                  const double *py, double *out);
void launch_neighbor_count(hipStream_t s, Device &d, int cur, int32_t *counts_orig);
void launch_neighbor_fill(hipStream_t s, Device &d, int cur, const int64_t *offsets_orig, int32_t *list,
                          int64_t *sort_key);
void launch_steady(hipStream_t s, Device &d, int cur);
void launch_download(hipStream_t s, Device &d, int cur);
void launch_download_rhs(hipStream_t s, Device &d);
void launch_upload(hipStream_t s, Device &d, const double *stage, const int32_t *ptype, const int32_t *inst_of,
                   int has_tem);

}  // namespace lssph
