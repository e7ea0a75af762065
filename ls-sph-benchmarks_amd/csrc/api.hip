// api.hip -- the C ABI of include/lssph.h: handle lifecycle, host<->device
// transfers, and the call sequences of the reference's RK2/RK4/PST/ghost
// subroutines expressed as kernel launches on the handle's stream.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace lssph;

struct lssph_handle {
  Device d;
  hipStream_t stream = nullptr;
  int device = -1;  // the device current at lssph_create (the stream's)
  std::vector<lssph_params_t> params;
  std::vector<InstDev> inst_host;
  std::vector<int32_t> n_int, n_ext;
  std::vector<int32_t> inst_of;  // upload index -> instance
  int cur = 0;          // records buffer holding the current state
  int passes = 0;       // mover passes since the last cell build
  bool spd_fresh = false;  // Device::spd holds the speeds of the current slots
  int max_passes = 1;   // passes a cell list serves (rk + 1 with a skin, else 1)
  double skin = 0.08;   // cell-grid skin as a fraction of h_eff
  int cell_split = 2;   // internal cell columns per stencil reach in x (InstDev::kx)
  bool uploaded = false, cells_ready = false;
  std::string err;
  // one captured step per starting state (records buffer, build buffers,
  // passes since the build)
  hipGraphExec_t graph[8] = {};
  int graph_end[8] = {};
  int reorder_interval = 50;
  int64_t steps_since_reorder = 0;
  std::vector<void *> allocs;
  // per-kernel event timing (eager calls only)
  bool profiling = false, capturing = false;
  std::vector<hipEvent_t> ev_pool;
  struct EvPair { hipEvent_t a, b; int group; };
  std::vector<EvPair> ev_pending;
  double kt_ms[5] = {0, 0, 0, 0, 0};
  int64_t kt_n[5] = {0, 0, 0, 0, 0};
};

static thread_local std::string g_create_error;

// ---- staging of lssph_probe_static ------------------------------------------
// Probe coordinates in, three results per probe out, for the entry point that
// calls no allocation routine.  Each GPU holds its own copy of the array for
// the life of the process; its address on a device is looked up once, with
// that device current, by the first lssph_create on it.  All calls on one GPU
// share it, so overlapping calls on one GPU race.
constexpr int kProbeStaticMax = 16384;  // probes per call (640 KiB of staging)
static __device__ __align__(256) double scratch[5 * kProbeStaticMax];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices] = {};

namespace {

int fail(lssph_handle *h, int code, const std::string &msg) {
  if (h)
    h->err = msg;
  else
    g_create_error = msg;
  return code;
}

#define HIP_TRY(h, expr)                                                               \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(h, LSSPH_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
void dfree(lssph_handle *h, T *&p) {
  if (!p) return;
  for (auto &a : h->allocs)
    if (a == (void *)p) {
      (void)hipFree(a);
      a = h->allocs.back();
      h->allocs.pop_back();
      break;
    }
  p = nullptr;
}

template <typename T>
int dalloc(lssph_handle *h, T **p, size_t count) {
  void *v = nullptr;
  hipError_t e = hipMalloc(&v, sizeof(T) * std::max<size_t>(count, 1));
  if (e != hipSuccess) return fail(h, LSSPH_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  h->allocs.push_back(v);
  *p = static_cast<T *>(v);
  return 0;
}

double kernel_norm(int kt, double h) {
  switch (kt) {
    case 1: return 5.0 / (14.0 * kPi * (h * h));
    case 2: return 7.0 / (478.0 * kPi * (h * h));
    case 3: return 7.0 / (4.0 * kPi * (h * h));
    case 4: return 9.0 / (4.0 * kPi * (h * h));
    default: return 39.0 / (14.0 * kPi * (h * h));
  }
}

int validate(const lssph_params_t *p, int n_inst, const int32_t *n_int, const int32_t *n_ext, std::string &msg) {
  const lssph_params_t &a = p[0];
  if (a.target_problem < 1 || a.target_problem > 4) return msg = "target_problem must be 1..4", 1;
  if (a.kernel_type < 1 || a.kernel_type > 5) return msg = "kernel_type must be 1..5", 1;
  if (a.sph_model < 1 || a.sph_model > 6) return msg = "sph_model must be 1..6", 1;
  if (a.wall_model < 1 || a.wall_model > 3) return msg = "wall_model must be 1..3", 1;
  if (a.rk != 2 && a.rk != 4) return msg = "rk must be 2 or 4", 1;
  if (a.precision != LSSPH_FP64 && a.precision != LSSPH_MIXED) return msg = "precision must be 0 or 1", 1;
  for (int k = 0; k < n_inst; ++k) {
    const lssph_params_t &b = p[k];
    if (b.target_problem != a.target_problem || b.kernel_type != a.kernel_type ||
        b.sph_model != a.sph_model || b.wall_model != a.wall_model || b.rk != a.rk ||
        b.periodic != a.periodic || b.precision != a.precision || b.u_bc_top != a.u_bc_top ||
        b.u_bc_bottom != a.u_bc_bottom || b.u_bc_left != a.u_bc_left || b.u_bc_right != a.u_bc_right ||
        b.tem_bc_top != a.tem_bc_top || b.tem_bc_bottom != a.tem_bc_bottom ||
        b.tem_bc_left != a.tem_bc_left || b.tem_bc_right != a.tem_bc_right)
      return msg = "model options differ between instances", 1;
    if (n_int[k] < 0 || n_ext[k] < 0) return msg = "negative particle count", 1;
    if (!(b.h > 0.0) || !(b.h_eff > 0.0) || !(b.dx > 0.0) || !(b.sp_mass > 0.0) || !(b.len_x > 0.0) ||
        !(b.len_y > 0.0))
      return msg = "h, h_eff, dx, sp_mass, len_x, len_y must be positive", 1;
    if (!(b.wl_thick >= 0.0)) return msg = "wl_thick must be >= 0", 1;
    // cells per direction are ints (cell width >= h_eff)
    if (!((b.len_x + 2.0 * b.wl_thick) / b.h_eff < 1e9) || !((b.len_y + 2.0 * b.wl_thick) / b.h_eff < 1e9))
      return msg = "problem too large for 32-bit indices: over 1e9 cells per direction", 1;
    if (b.periodic && n_ext[k] != 0) return msg = "a periodic instance has no wall particles", 1;
    if (b.periodic && (std::floor(b.len_x / b.h_eff) < 3 || std::floor(b.len_y / b.h_eff) < 3))
      return msg = "a periodic box needs at least 3 cells of width >= h_eff per direction", 1;
  }
  return 0;
}

// Internal cell width: h_eff + skin * h_eff, reduced for periodic boxes so
// that at least 3 cells fit per direction (skin >= 0).
double cell_width(const lssph_params_t &p, double skin) {
  double w = p.h_eff * (1.0 + skin);
  if (p.periodic) w = std::min(w, std::min(p.len_x, p.len_y) / 3.0);
  return std::max(w, p.h_eff);
}

InstDev make_inst(const lssph_params_t &p, double skin, int kx) {
  InstDev I{};
  I.h = p.h;
  I.inv_h = 1.0 / p.h;
  I.inv_h2 = I.inv_h / p.h;
  I.inv_h_sq = I.inv_h * I.inv_h;
  I.inv_h3 = I.inv_h2 / p.h;
  I.inv_h4 = I.inv_h3 / p.h;
  I.h_eff = p.h_eff;
  I.h_eff_sq = p.h_eff * p.h_eff;
  I.inv_h_eff = 1.0 / p.h_eff;
  I.threshold = p.threshold;
  I.sp_mass = p.sp_mass;
  I.w_ave = p.w_ave;
  I.wl_thick = p.wl_thick;
  I.len_x = p.len_x;
  I.len_y = p.len_y;
  I.w_norm = kernel_norm(p.kernel_type, p.h);
  I.dw_fac = p.kernel_type == 1 ? -3.0 / p.h : (p.kernel_type == 2 ? -5.0 / p.h : 1.0 / p.h);
  I.rho_ref = p.rho_ref;
  I.vis_ref = p.vis_ref;
  I.c_sq = p.sound_speed * p.sound_speed;
  I.thermal_dif = p.thermal_dif;
  I.alpha_ref = p.alpha_ref;
  I.gravity = p.gravity;
  I.tem_ave = p.tem_ave;
  I.zeta_sq = p.zeta_rsst * p.zeta_rsst;
  I.xi_sq = p.xi_vim * p.xi_vim;
  I.delta = p.coe_delta_sph * p.sound_speed * p.h * p.zeta_rsst;
  I.pst_c = p.pst_c;
  I.dt = p.dt;
  I.dt_half = p.dt / 2.0;
  I.dt_third = p.dt / 3.0;
  I.dt_sixth = p.dt / 6.0;
  I.u_top = p.u_top;
  I.tem_top = p.tem_top;
  I.tem_bottom = p.tem_bottom;
  I.tem_left = p.tem_left;
  I.tem_right = p.tem_right;
  I.diff_coef = p.diff_coef;
  I.diff_source = p.diff_source;
  I.cond_max = p.cond_max > 0.0 ? p.cond_max : 1.0e4;
  I.x_min = 0.0 + p.wl_thick;
  I.x_max = p.len_x + p.wl_thick;
  I.y_min = 0.0 + p.wl_thick;
  I.y_max = p.len_y + p.wl_thick;
  I.periodic = p.periodic;
  const double w = cell_width(p, skin);
  double sk = w - p.h_eff;  // the skin this instance gets
  // cells are w tall and w / kx wide (kx columns per stencil reach)
  I.kx = kx;
  if (p.periodic) {
    I.nx = (int)std::min(kx * std::floor(p.len_x / w), 2e9);  // > 2^30 cells fails setup_grid's check
    I.ny = (int)std::floor(p.len_y / w);
    I.inv_cw_x = (double)I.nx / p.len_x;
    I.inv_cw_y = (double)I.ny / p.len_y;
    I.rnx = (int)std::floor(p.len_x / p.h_eff);
    I.rny = (int)std::floor(p.len_y / p.h_eff);
    I.inv_rcw_x = (double)I.rnx / p.len_x;
    I.inv_rcw_y = (double)I.rny / p.len_y;
    // whole cells per period: a stencil reaches at least w
    sk = std::min(p.len_x / (I.nx / kx), p.len_y / I.ny) - p.h_eff;
  } else {
    I.nx = (int)std::min(std::floor((p.len_x + 2.0 * p.wl_thick) * kx / w) + 1.0, 2e9);
    I.ny = (int)std::floor((p.len_y + 2.0 * p.wl_thick) / w) + 1;
    I.inv_cw_x = kx / w;
    I.inv_cw_y = 1.0 / w;
    // init_cell_mod.f90:36-41
    I.rnx = (int)std::ceil((p.len_x + 2.0 * p.wl_thick) / p.h_eff);
    I.rny = (int)std::ceil((p.len_y + 2.0 * p.wl_thick) / p.h_eff);
  }
  // each particle may move 0.45 skin between builds, which leaves a tenth
  // of the skin as margin for rounding in the binning; skin 0 asks for a
  // build after every pass whatever width the cells have
  I.skin_half_sq = skin > 0.0 ? (0.45 * sk) * (0.45 * sk) : 0.0;
  return I;
}

int check_counters(lssph_handle *h) {
  int32_t c[kNumCounters];
  HIP_TRY(h, hipMemcpyAsync(c, h->d.counters, sizeof c, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (c[kEscaped] > 0) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "%d particle(s) or marker(s) left the cell grid", c[kEscaped]);
    return fail(h, LSSPH_E_DOMAIN, buf);
  }
  if (c[kSkinExceeded] > 0) {
    return fail(h, LSSPH_E_DOMAIN,
                "a particle moved farther than half the cell-list skin between cell builds (lssph_set_skin "
                "raises the skin; 0 rebuilds the cells after every pass)");
  }
  if (c[kWaveListOverflow] > 0) {
    char buf[200];
    std::snprintf(buf, sizeof buf,
                  "%d particle(s) have more than %d neighbors, the limit of the SPH_MODEL 6 kernels",
                  c[kWaveListOverflow], kWaveList);
    return fail(h, LSSPH_E_CAPACITY, buf);
  }
  if (c[kStrayOverflow] > 0) {
    char buf[200];
    std::snprintf(buf, sizeof buf,
                  "more than %d particles moved farther than one cell in one stage (%d over the limit); "
                  "the time step is far above its stability limit",
                  kStrayCap, c[kStrayOverflow]);
    return fail(h, LSSPH_E_DOMAIN, buf);
  }
  return 0;
}

bool is_diffusion(const lssph_handle *h) { return h->d.opt.target_problem == 1; }

enum { kGrpLsa = 0, kGrpPst, kGrpCells, kGrpGhost, kGrpOther };

hipEvent_t take_event(lssph_handle *h) {
  if (!h->ev_pool.empty()) {
    hipEvent_t e = h->ev_pool.back();
    h->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets the launches made by f() with an event pair when profiling.
template <typename F>
void timed(lssph_handle *h, int group, F &&f) {
  if (!h->profiling || h->capturing) {
    f();
    return;
  }
  hipEvent_t a = take_event(h), b = take_event(h);
  if (!a || !b) {
    f();
    return;
  }
  (void)hipEventRecord(a, h->stream);
  f();
  (void)hipEventRecord(b, h->stream);
  h->ev_pending.push_back({a, b, group});
}

// A mover pass (RK stage or shift) moves every fluid particle.  The
// reference rebuilds its cell list after each (update_cell,
// RK2_mod.f90:116, RK4_mod.f90:295, PST_mod.f90:101); here the cells are
// h_eff + skin wide and one list serves max_passes passes -- a time step --
// with every move checked against the skin, and every pass still tests
// r^2 <= h_eff^2 at the current positions.  The pass that uses a list up
// also writes the new cell keys, and the incremental build follows it.  The
// shift (up to 0.2 h per particle) always ends a list.
// Returns whether a cell build followed the pass.
template <typename F>
bool mover_pass(lssph_handle *h, int group, bool ends_list, F &&launch) {
  Device &d = h->d;
  const bool moving = !is_diffusion(h);
  const bool build = moving && (ends_list || h->passes + 1 >= h->max_passes);
  const int guard = (moving && !build && h->max_passes > 1) ? 1 : 0;
  timed(h, group, [&] { launch(guard, build ? 1 : 0); });
  h->cur ^= 1;
  h->spd_fresh = false;
  if (!moving) return false;
  ++h->passes;
  if (build) {
    timed(h, kGrpCells, [&] { launch_rebuild(h->stream, d, h->cur, h->cur ^ 1); });
    h->cur ^= 1;
    h->passes = 0;
  }
  return build;
}

// One RK stage pass: fused operator + update, the cell list, and ghost on
// the new state (RK2_mod.f90:67-126, RK4_mod.f90:205-306).  The diffusion
// test's particles do not move (RK4_mod.f90:295-297).
void enqueue_stage(lssph_handle *h, StageMode m) {
  Device &d = h->d;
  // the last stage of a step that shifts stores the speeds U_max needs
  const bool spd = (m == kRk2S2 || m == kRk4S4) && !is_diffusion(h);
  const bool built = mover_pass(h, kGrpLsa, false, [&](int guard, int keys) {
    launch_lsa(h->stream, d, h->cur, m, guard, keys, spd);
  });
  h->spd_fresh = spd && !built;  // a build moves the slots
  if (d.n_ext) timed(h, kGrpGhost, [&] { launch_ghost(h->stream, d, h->cur); });
}

void enqueue_rk(lssph_handle *h) {
  // isolation mode: the device's step number, which a retirement records
  if (h->d.retired) launch_add_i32(h->stream, h->d.counters + kStepNo, 1);
  if (h->d.opt.rk == 2) {
    enqueue_stage(h, kRk2S1);
    enqueue_stage(h, kRk2S2);
  } else {
    enqueue_stage(h, kRk4S1);
    enqueue_stage(h, kRk4S2);
    enqueue_stage(h, kRk4S3);
    enqueue_stage(h, kRk4S4);
  }
}

// PST_mod.f90:37-114
void enqueue_pst(lssph_handle *h) {
  Device &d = h->d;
  timed(h, kGrpOther, [&] {
    if (h->spd_fresh)
      launch_umax_spd(h->stream, d);
    else
      launch_umax(h->stream, d, h->cur);
  });
  mover_pass(h, kGrpPst, true, [&](int guard, int keys) { launch_pst(h->stream, d, h->cur, guard, keys); });
  if (d.n_ext) timed(h, kGrpGhost, [&] { launch_ghost(h->stream, d, h->cur); });
}

void enqueue_step(lssph_handle *h) {
  enqueue_rk(h);
  if (!is_diffusion(h)) enqueue_pst(h);
}

// Physical re-sort: permutes the per-particle arrays indexed by physical id
// into the current slot order, so that the RK accumulators a slot touches
// stay contiguous; slots then map to themselves.
int reorder(lssph_handle *h) {
  Device &d = h->d;
  hipStream_t s = h->stream;
  const int n = d.n_total;
  launch_gather_i32(s, n, d.items, d.info, d.stage_i32);
  HIP_TRY(h, hipMemcpyAsync(d.info, d.stage_i32, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
  launch_gather_i32(s, n, d.items, d.key, d.stage_i32);
  HIP_TRY(h, hipMemcpyAsync(d.key, d.stage_i32, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
  launch_gather_i32(s, n, d.items, d.orig_of, d.stage_i32);
  HIP_TRY(h, hipMemcpyAsync(d.orig_of, d.stage_i32, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
  launch_gather_f64(s, n, d.items, d.f_prev, d.stage_f64);
  HIP_TRY(h, hipMemcpyAsync(d.f_prev, d.stage_f64, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  launch_inverse_perm(s, n, d.orig_of, d.phys_of);
  launch_iota(s, d.items, n);
  launch_iota(s, d.slot_of, n);
  launch_vm_init(s, d, h->cur);
  h->steps_since_reorder = 0;
  h->spd_fresh = false;
  return 0;
}

// Full cell build of the current records (new buffers become current).
void full_build(lssph_handle *h, int fixed_key_mode) {
  launch_bin(h->stream, h->d, h->cur, h->cur ^ 1, fixed_key_mode);
  h->cur ^= 1;
  h->passes = 0;
  h->spd_fresh = false;
}

void drop_graphs(lssph_handle *h) {
  for (int i = 0; i < 8; ++i)
    if (h->graph[i]) {
      (void)hipGraphExecDestroy(h->graph[i]);
      h->graph[i] = nullptr;
    }
}

// Per-instance constants and the internal cell grid for the handle's skin;
// returns the total cell count.  Passes per cell list: rk + 1 with a skin
// (one build per time step), 1 without.
int64_t setup_grid(lssph_handle *h) {
  h->inst_host.clear();
  int64_t nc = 0;
  bool all_skin = true;
  for (size_t k = 0; k < h->params.size(); ++k) {
    InstDev I = make_inst(h->params[k], h->skin, h->cell_split);
    I.cell_off = (int32_t)nc;
    nc += (int64_t)I.nx * I.ny;
    all_skin = all_skin && I.skin_half_sq > 0.0;
    h->inst_host.push_back(I);
  }
  h->d.n_cells = (int)std::min<int64_t>(nc, 1ll << 30);
  h->max_passes = all_skin ? h->params[0].rk + 1 : 1;
  return nc;
}

// Arrays sized by the cell grid; uploads the instance constants and the
// cell -> instance map.
int alloc_grid(lssph_handle *h) {
  Device &d = h->d;
  dfree(h, d.cell_count);
  dfree(h, d.cstart_buf[0]);
  dfree(h, d.cstart_buf[1]);
  dfree(h, d.cell_inst);
  dfree(h, d.narr);
  dfree(h, d.arr);
  dfree(h, d.scan_tmp);
  int rc = 0;
  rc = rc ? rc : dalloc(h, &d.cell_count, d.n_cells);
  for (int b = 0; b < 2 && !rc; ++b) rc = dalloc(h, &d.cstart_buf[b], d.n_cells + 1);
  rc = rc ? rc : dalloc(h, &d.cell_inst, d.n_cells);
  rc = rc ? rc : dalloc(h, &d.narr, d.n_cells);
  rc = rc ? rc : dalloc(h, &d.arr, (size_t)kArrivals * d.n_cells);
  rc = rc ? rc : dalloc(h, &d.scan_tmp, scan_tile_count(d.n_cells) + 1);
  if (rc) return rc;
  HIP_TRY(h, hipMemcpy(d.inst, h->inst_host.data(), sizeof(InstDev) * d.n_inst, hipMemcpyHostToDevice));
  std::vector<int32_t> ci(d.n_cells);
  for (int k = 0; k < d.n_inst; ++k) {
    const InstDev &I = h->inst_host[k];
    for (int c = 0; c < I.nx * I.ny; ++c) ci[I.cell_off + c] = k;
  }
  HIP_TRY(h, hipMemcpy(d.cell_inst, ci.data(), sizeof(int32_t) * d.n_cells, hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemset(d.cell_count, 0, sizeof(int32_t) * d.n_cells));
  select_build(d, d.ib);
  return 0;
}

// Device address of the staging array on the handle's device, looked up once
// per device (the array's host-side name is not a device address).
int resolve_scratch(lssph_handle *h) {
  HIP_TRY(h, hipGetDevice(&h->device));
  if (h->device < 0 || h->device >= kMaxDevices) return fail(h, LSSPH_E_HIP, "device index beyond 64");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[h->device]) return 0;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(h, LSSPH_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[h->device] = static_cast<double *>(p);
  return 0;
}

int ready(lssph_handle *h) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  if (!h->uploaded) return fail(h, LSSPH_E_STATE, "call lssph_upload first");
  if (!h->cells_ready) return fail(h, LSSPH_E_STATE, "call lssph_init_cell first");
  return 0;
}

}  // namespace

extern "C" {

int32_t lssph_create(const lssph_params_t *params, int32_t n_inst, const int32_t *n_int, const int32_t *n_ext,
                     lssph_handle_t *out) {
  if (!params || !n_int || !n_ext || !out || n_inst < 1)
    return fail(nullptr, LSSPH_E_ARG, "lssph_create: null argument or n_inst < 1");
  *out = nullptr;
  std::string msg;
  if (validate(params, n_inst, n_int, n_ext, msg)) return fail(nullptr, LSSPH_E_ARG, "lssph_create: " + msg);
  auto *h = new lssph_handle();
  h->params.assign(params, params + n_inst);
  h->n_int.assign(n_int, n_int + n_inst);
  h->n_ext.assign(n_ext, n_ext + n_inst);
  Device &d = h->d;
  d.n_inst = n_inst;
  int64_t ni = 0, ne = 0;
  for (int k = 0; k < n_inst; ++k) {
    ni += n_int[k];
    ne += n_ext[k];
  }
  const int64_t nc = setup_grid(h);
  if (ni + ne > kMaxParticles) {
    delete h;
    return fail(nullptr, LSSPH_E_ARG,
                "lssph_create: more than " + std::to_string(kMaxParticles) +
                    " particles in one handle (neighbor-list entries hold 27-bit slots); split the problem "
                    "over several handles");
  }
  if (nc > (1ll << 30) || n_inst >= (1 << 23)) {
    delete h;
    return fail(nullptr, LSSPH_E_ARG, "lssph_create: problem too large for 32-bit indices");
  }
  d.n_int = (int)ni;
  d.n_ext = (int)ne;
  d.n_total = (int)(ni + ne);
  const lssph_params_t &a = params[0];
  d.opt = Options{a.target_problem, a.kernel_type, a.wall_model, a.rk, a.u_bc_top, a.u_bc_bottom,
                  a.u_bc_left, a.u_bc_right, a.tem_bc_top, a.tem_bc_bottom, a.tem_bc_left, a.tem_bc_right,
                  a.target_problem == 4 ? 1 : 0, a.precision, 0, a.sph_model};
  for (int k = 0; k < n_inst; ++k)
    if (params[k].coe_delta_sph != 0.0) d.opt.delta = 1;
  {
    int64_t nmax = 1;
    for (int k = 0; k < n_inst; ++k) nmax = std::max<int64_t>(nmax, (int64_t)n_int[k] + n_ext[k]);
    d.red_blocks = (int)std::min<int64_t>(1024, (nmax + 4095) / 4096);
  }
  {
    // on-chip list length from the expected neighbor count pi (h_eff/dx)^2
    // with a quarter of headroom for disorder (30 -> 39 entries at the
    // default h_eff = 3.1 dx, 68 -> 86 for the quintic spline's 3 h), within
    // the 64 KiB of LDS of a block
    double nmax = 0.0;
    for (int k = 0; k < n_inst; ++k) {
      const double r = params[k].h_eff / params[k].dx;
      nmax = std::max(nmax, kPi * r * r);
    }
    const int want = (int)std::ceil(1.25 * nmax) + 1;
    const int cap = 64 * 1024 / (4 * d.block) - 1;
    d.list_len = std::min(std::max(d.list_len, want), std::min(cap, 128));
  }
  // upload index -> instance: fluid blocks, then wall blocks, in instance order
  h->inst_of.resize(d.n_total);
  {
    int64_t f = 0, w = ni;
    for (int k = 0; k < n_inst; ++k) {
      for (int i = 0; i < n_int[k]; ++i) h->inst_of[f++] = k;
      for (int i = 0; i < n_ext[k]; ++i) h->inst_of[w++] = k;
    }
  }
  const size_t n = d.n_total;
  int rc = 0;
  rc = rc ? rc : dalloc(h, &d.inst, n_inst);
  rc = rc ? rc : dalloc(h, &d.mir[0], n);
  rc = rc ? rc : dalloc(h, &d.mir[1], n);
  if (mixed_path(d))
    for (int b = 0; b < 2 && !rc; ++b) rc = dalloc(h, &d.nbf[b], n);
  rc = rc ? rc : dalloc(h, &d.xbuild, 2 * n);
  rc = rc ? rc : dalloc(h, &d.info, n);
  rc = rc ? rc : dalloc(h, &d.key, n);
  rc = rc ? rc : dalloc(h, &d.orig_of, n);
  rc = rc ? rc : dalloc(h, &d.phys_of, n);
  for (int b = 0; b < 2 && !rc; ++b) rc = dalloc(h, &d.items_buf[b], n);
  rc = rc ? rc : dalloc(h, &d.kslot, n);
  rc = rc ? rc : dalloc(h, &d.strays, 2 * kStrayCap);
  rc = rc ? rc : dalloc(h, &d.nstray, 1);
  rc = rc ? rc : dalloc(h, &d.sorig, n);
  rc = rc ? rc : dalloc(h, &d.sinfo, n);
  rc = rc ? rc : dalloc(h, &d.slot_of, n);
  rc = rc ? rc : dalloc(h, &d.fb_list, n);
  rc = rc ? rc : dalloc(h, &d.perm_tmp, n);
  for (int f = 0; f < 6 && !rc; ++f) rc = dalloc(h, &d.old_[f], n);
  for (int f = 0; f < 6 && !rc; ++f) rc = dalloc(h, &d.new_[f], n);
  for (int f = 0; f < 4 && !rc; ++f) rc = dalloc(h, &d.rhs[f], n);
  rc = rc ? rc : dalloc(h, &d.f_prev, n);
  rc = rc ? rc : dalloc(h, &d.spd, n);
  const size_t m = d.n_ext;
  rc = rc ? rc : dalloc(h, &d.vm_x, m);
  rc = rc ? rc : dalloc(h, &d.vm_y, m);
  rc = rc ? rc : dalloc(h, &d.vm_nx, m);
  rc = rc ? rc : dalloc(h, &d.vm_ny, m);
  rc = rc ? rc : dalloc(h, &d.vm_u, m);
  rc = rc ? rc : dalloc(h, &d.vm_v, m);
  rc = rc ? rc : dalloc(h, &d.vm_pre, m);
  rc = rc ? rc : dalloc(h, &d.vm_tem, m);
  rc = rc ? rc : dalloc(h, &d.vm_ptype, m);
  rc = rc ? rc : dalloc(h, &d.vm_inst, m);
  rc = rc ? rc : dalloc(h, &d.vm_cell, m);
  rc = rc ? rc : dalloc(h, &d.vm_wall, m);
  rc = rc ? rc : dalloc(h, &d.umax, n_inst);
  rc = rc ? rc : dalloc(h, &d.dfmax, n_inst);
  rc = rc ? rc : dalloc(h, &d.steady, n_inst);
  rc = rc ? rc : dalloc(h, &d.counters, kNumCounters);
  rc = rc ? rc : dalloc(h, &d.retired_buf, 2 * n_inst);
  rc = rc ? rc : dalloc(h, &d.stats64, kNumStats64);
  rc = rc ? rc : dalloc(h, &d.stage_f64, 8 * n);
  rc = rc ? rc : dalloc(h, &d.stage_i32, n);
  if (!rc) {
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = fail(h, LSSPH_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  rc = rc ? rc : resolve_scratch(h);
  rc = rc ? rc : alloc_grid(h);
  if (rc) {
    g_create_error = h->err;
    lssph_destroy(h);
    return rc;
  }
  select_build(d, 0);
  // both builds' slot arrays start with valid particle ids
  for (int b = 0; b < 2; ++b) launch_fill_i32(h->stream, d.items_buf[b], 0, d.n_total);
  launch_fill_i32(h->stream, d.counters, 0, kNumCounters);
  launch_fill_i32(h->stream, d.nstray, 0, 1);
  launch_fill_u64(h->stream, d.stats64, 0ull, kNumStats64);
  *out = h;
  return 0;
}

int32_t lssph_destroy(lssph_handle_t h) {
  if (!h) return 0;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graphs(h);
  for (auto &e : h->ev_pending) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  for (auto e : h->ev_pool) (void)hipEventDestroy(e);
  for (void *p : h->allocs) (void)hipFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

const char *lssph_last_error(lssph_handle_t h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int32_t lssph_upload(lssph_handle_t h, const int32_t *ptype, const double *x, const double *y, const double *u,
                     const double *v, const double *rho, const double *pre, const double *tem) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  if (!ptype || !x || !y || !u || !v || !rho || !pre) return fail(h, LSSPH_E_ARG, "lssph_upload: null array");
  Device &d = h->d;
  const int has_tem = d.opt.has_tem;
  if (has_tem && !tem) return fail(h, LSSPH_E_ARG, "lssph_upload: tem is required for target_problem 4");
  const size_t n = d.n_total;
  for (size_t i = 0; i < n; ++i) {
    const bool fluid_slot = (int64_t)i < d.n_int;
    if ((ptype[i] == 0) != fluid_slot || ptype[i] < 0 || ptype[i] > 8)
      return fail(h, LSSPH_E_ARG, "lssph_upload: ptype must be 0 for the fluid block and 1..8 for walls");
    if (!(rho[i] > 0.0)) return fail(h, LSSPH_E_ARG, "lssph_upload: rho must be positive");
  }
  hipStream_t s = h->stream;
  drop_graphs(h);
  const double *src[7] = {x, y, u, v, rho, pre, tem};
  for (int f = 0; f < 7; ++f)
    if (src[f]) HIP_TRY(h, hipMemcpyAsync(d.stage_f64 + f * n, src[f], sizeof(double) * n, hipMemcpyHostToDevice, s));
  HIP_TRY(h, hipMemcpyAsync(d.stage_i32, ptype, sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  HIP_TRY(h, hipMemcpyAsync(d.perm_tmp, h->inst_of.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  launch_upload(s, d, d.stage_f64, d.stage_i32, d.perm_tmp, has_tem);
  // identity permutation (records in upload order in buffer 0), then the
  // initial (cell, index) sort
  launch_iota(s, d.orig_of, (int)n);
  launch_iota(s, d.items, (int)n);
  launch_iota(s, d.slot_of, (int)n);
  launch_fill_i32(s, d.counters, 0, kNumCounters);
  launch_fill_i32(s, d.retired_buf, 0, 2 * d.n_inst);
  HIP_TRY(h, hipMemcpyAsync(d.inst, h->inst_host.data(), sizeof(InstDev) * d.n_inst, hipMemcpyHostToDevice, s));
  launch_fill_i32(s, d.cell_count, 0, d.n_cells);
  h->cur = 0;
  full_build(h, 1);
  // f_prev = 1 (setup_closed_box_mod.f90:180-183); permuted along with the rest
  {
    std::vector<double> ones(n, 1.0);
    HIP_TRY(h, hipMemcpyAsync(d.f_prev, ones.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(h, hipStreamSynchronize(s));
  }
  launch_inverse_perm(s, d.n_total, d.orig_of, d.phys_of);
  HIP_TRY(h, hipStreamSynchronize(s));
  int rc = reorder(h);
  if (rc) return rc;
  h->uploaded = true;
  h->cells_ready = false;
  rc = check_counters(h);
  if (rc) return rc;
  HIP_TRY(h, hipGetLastError());
  return 0;
}

int32_t lssph_download(lssph_handle_t h, int32_t *ptype, double *x, double *y, double *u, double *v, double *rho,
                       double *pre, double *tem) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  if (!h->uploaded) return fail(h, LSSPH_E_STATE, "call lssph_upload first");
  Device &d = h->d;
  hipStream_t s = h->stream;
  const size_t n = d.n_total;
  launch_download(s, d, h->cur);
  double *dst[7] = {x, y, u, v, rho, pre, d.opt.has_tem ? tem : nullptr};
  for (int f = 0; f < 7; ++f)
    if (dst[f]) HIP_TRY(h, hipMemcpyAsync(dst[f], d.stage_f64 + f * n, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  if (ptype) HIP_TRY(h, hipMemcpyAsync(ptype, d.stage_i32, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  HIP_TRY(h, hipStreamSynchronize(s));
  HIP_TRY(h, hipGetLastError());
  return check_counters(h);
}

int32_t lssph_download_cell_index(lssph_handle_t h, int32_t *idx_sp) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  if (!idx_sp) return fail(h, LSSPH_E_ARG, "lssph_download_cell_index: null array");
  if (!h->uploaded) return fail(h, LSSPH_E_STATE, "call lssph_upload first");
  Device &d = h->d;
  launch_download(h->stream, d, h->cur);
  HIP_TRY(h, hipMemcpyAsync(idx_sp, d.perm_tmp, sizeof(int32_t) * d.n_total, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return 0;
}

int32_t lssph_init_cell(lssph_handle_t h) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  if (!h->uploaded) return fail(h, LSSPH_E_STATE, "call lssph_upload first");
  full_build(h, 0);
  h->cells_ready = true;
  HIP_TRY(h, hipGetLastError());
  return 0;
}

int32_t lssph_update_cell(lssph_handle_t h) {
  if (int rc = ready(h)) return rc;
  full_build(h, 0);
  HIP_TRY(h, hipGetLastError());
  return 0;
}

int32_t lssph_ghost(lssph_handle_t h) {
  if (int rc = ready(h)) return rc;
  launch_ghost(h->stream, h->d, h->cur);
  HIP_TRY(h, hipGetLastError());
  return 0;
}

int32_t lssph_calculate_rhs(lssph_handle_t h) {
  if (int rc = ready(h)) return rc;
  launch_lsa(h->stream, h->d, h->cur, kRhsOnly, 0, 0);
  HIP_TRY(h, hipGetLastError());
  return 0;
}

int32_t lssph_download_rhs(lssph_handle_t h, double *rhs_u, double *rhs_v, double *rhs_rho, double *rhs_tem) {
  if (int rc = ready(h)) return rc;
  Device &d = h->d;
  launch_download_rhs(h->stream, d);
  double *dst[4] = {rhs_u, rhs_v, rhs_rho, d.opt.has_tem ? rhs_tem : nullptr};
  for (int f = 0; f < 4; ++f)
    if (dst[f])
      HIP_TRY(h, hipMemcpyAsync(dst[f], d.stage_f64 + (size_t)f * d.n_int, sizeof(double) * d.n_int,
                                hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return 0;
}

// The periodic physical re-sort at the start of a time step (no RK
// accumulator holds state there); both the eager RK calls and lssph_step
// count their steps toward it.
static int maybe_reorder(lssph_handle *h) {
  if (h->reorder_interval > 0 && !is_diffusion(h) && h->steps_since_reorder >= h->reorder_interval)
    return reorder(h);
  return 0;
}

int32_t lssph_rk2(lssph_handle_t h) {
  if (int rc = ready(h)) return rc;
  if (h->d.opt.rk != 2) return fail(h, LSSPH_E_ARG, "handle was created with rk = 4");
  if (int rc = maybe_reorder(h)) return rc;
  enqueue_rk(h);
  ++h->steps_since_reorder;
  HIP_TRY(h, hipGetLastError());
  return 0;
}

int32_t lssph_rk4(lssph_handle_t h) {
  if (int rc = ready(h)) return rc;
  if (h->d.opt.rk != 4) return fail(h, LSSPH_E_ARG, "handle was created with rk = 2");
  if (int rc = maybe_reorder(h)) return rc;
  enqueue_rk(h);
  ++h->steps_since_reorder;
  HIP_TRY(h, hipGetLastError());
  return 0;
}

int32_t lssph_pst(lssph_handle_t h) {
  if (int rc = ready(h)) return rc;
  enqueue_pst(h);
  HIP_TRY(h, hipGetLastError());
  return 0;
}

int32_t lssph_step(lssph_handle_t h, int32_t nsteps) {
  if (int rc = ready(h)) return rc;
  if (nsteps < 0) return fail(h, LSSPH_E_ARG, "nsteps must be >= 0");
  for (int32_t i = 0; i < nsteps; ++i) {
    if (int rc = maybe_reorder(h)) return rc;
    // the captured step depends on the records buffer, the build buffers and
    // whether the cell list is fresh
    if (h->passes != 0) full_build(h, 0);
    const int c0 = h->cur | (h->d.ib << 1);
    if (!h->graph[c0]) {
      hipGraph_t g = nullptr;
      HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      h->capturing = true;
      enqueue_step(h);
      h->capturing = false;
      if (h->passes != 0) full_build(h, 0);  // a list that would outlive the step
      const int c1 = h->cur | (h->d.ib << 1);
      h->cur = c0 & 1;
      select_build(h->d, (c0 >> 1) & 1);
      HIP_TRY(h, hipStreamEndCapture(h->stream, &g));
      hipError_t e = hipGraphInstantiate(&h->graph[c0], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        h->graph[c0] = nullptr;
        return fail(h, LSSPH_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
      h->graph_end[c0] = c1;
    }
    HIP_TRY(h, hipGraphLaunch(h->graph[c0], h->stream));
    h->cur = h->graph_end[c0] & 1;
    select_build(h->d, (h->graph_end[c0] >> 1) & 1);
    h->passes = 0;
    ++h->steps_since_reorder;
  }
  return 0;
}

int32_t lssph_check_steady_state(lssph_handle_t h, int32_t *is_steady, double *df_max) {
  if (int rc = ready(h)) return rc;
  Device &d = h->d;
  launch_steady(h->stream, d, h->cur);
  std::vector<unsigned long long> bits(d.n_inst);
  std::vector<int32_t> st(d.n_inst);
  HIP_TRY(h, hipMemcpyAsync(bits.data(), d.dfmax, sizeof(unsigned long long) * d.n_inst, hipMemcpyDeviceToHost,
                            h->stream));
  HIP_TRY(h, hipMemcpyAsync(st.data(), d.steady, sizeof(int32_t) * d.n_inst, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  for (int k = 0; k < d.n_inst; ++k) {
    double v;
    std::memcpy(&v, &bits[k], sizeof v);
    if (df_max) df_max[k] = v;
    if (is_steady) is_steady[k] = st[k];
  }
  return 0;
}

int32_t lssph_download_vm(lssph_handle_t h, int32_t *ptype, double *x, double *y, double *u, double *v,
                          double *pre, double *tem) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  if (!h->uploaded) return fail(h, LSSPH_E_STATE, "call lssph_upload first");
  Device &d = h->d;
  const size_t m = d.n_ext;
  hipStream_t s = h->stream;
  if (m == 0) return 0;
  if (ptype) HIP_TRY(h, hipMemcpyAsync(ptype, d.vm_ptype, sizeof(int32_t) * m, hipMemcpyDeviceToHost, s));
  const double *src[6] = {d.vm_x, d.vm_y, d.vm_u, d.vm_v, d.vm_pre, d.vm_tem};
  double *dst[6] = {x, y, u, v, pre, d.opt.has_tem ? tem : nullptr};
  for (int f = 0; f < 6; ++f)
    if (dst[f]) HIP_TRY(h, hipMemcpyAsync(dst[f], src[f], sizeof(double) * m, hipMemcpyDeviceToHost, s));
  HIP_TRY(h, hipStreamSynchronize(s));
  return 0;
}

int32_t lssph_neighbors(lssph_handle_t h, int32_t *counts, int32_t *list, int64_t capacity, int64_t *total) {
  if (int rc = ready(h)) return rc;
  if (!counts) return fail(h, LSSPH_E_ARG, "lssph_neighbors: null counts");
  Device &d = h->d;
  hipStream_t s = h->stream;
  int32_t *dcounts = nullptr;
  int64_t *doff = nullptr, *dkey = nullptr;
  int32_t *dlist = nullptr;
  HIP_TRY(h, hipMalloc(&dcounts, sizeof(int32_t) * std::max(d.n_int, 1)));
  launch_neighbor_count(s, d, h->cur, dcounts);
  hipError_t e = hipMemcpyAsync(counts, dcounts, sizeof(int32_t) * d.n_int, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(dcounts);
    return fail(h, LSSPH_E_HIP, std::string("lssph_neighbors: ") + hipGetErrorString(e));
  }
  std::vector<int64_t> off(d.n_int + 1, 0);
  for (int i = 0; i < d.n_int; ++i) off[i + 1] = off[i] + counts[i];
  const int64_t tot = off[d.n_int];
  if (total) *total = tot;
  int rc = 0;
  if (tot > capacity || !list) {
    rc = fail(h, LSSPH_E_CAPACITY, "lssph_neighbors: list capacity too small");
  } else if (tot > 0) {
    if (hipMalloc(&doff, sizeof(int64_t) * (d.n_int + 1)) != hipSuccess ||
        hipMalloc(&dlist, sizeof(int32_t) * tot) != hipSuccess ||
        hipMalloc(&dkey, sizeof(int64_t) * tot) != hipSuccess) {
      rc = fail(h, LSSPH_E_NOMEM, "lssph_neighbors: hipMalloc failed");
    } else {
      (void)hipMemcpyAsync(doff, off.data(), sizeof(int64_t) * (d.n_int + 1), hipMemcpyHostToDevice, s);
      launch_neighbor_fill(s, d, h->cur, doff, dlist, dkey);
      (void)hipMemcpyAsync(list, dlist, sizeof(int32_t) * tot, hipMemcpyDeviceToHost, s);
      if (hipStreamSynchronize(s) != hipSuccess) rc = fail(h, LSSPH_E_HIP, "lssph_neighbors: kernel failed");
    }
  }
  (void)hipFree(dcounts);
  if (doff) (void)hipFree(doff);
  if (dlist) (void)hipFree(dlist);
  if (dkey) (void)hipFree(dkey);
  return rc;
}

int32_t lssph_pair_lists(lssph_handle_t h, int32_t cap, int32_t *counts, int32_t *list, int32_t *path) {
  if (int rc = ready(h)) return rc;
  if (!counts || !list || !path || cap < 1) return fail(h, LSSPH_E_ARG, "lssph_pair_lists: null array or cap < 1");
  if (h->d.opt.sph_model == 6)
    return fail(h, LSSPH_E_ARG, "lssph_pair_lists: the SPH_MODEL 6 kernels gather per lane group (lssph_neighbors)");
  Device &d = h->d;
  hipStream_t s = h->stream;
  const size_t ni = std::max(d.n_int, 1);
  int32_t *dc = nullptr, *dl = nullptr, *dp = nullptr;
  int rc = 0;
  if (hipMalloc(&dc, sizeof(int32_t) * ni) != hipSuccess || hipMalloc(&dp, sizeof(int32_t) * ni) != hipSuccess ||
      hipMalloc(&dl, sizeof(int32_t) * ni * (size_t)cap) != hipSuccess) {
    rc = fail(h, LSSPH_E_NOMEM, "lssph_pair_lists: hipMalloc failed");
  } else {
    (void)hipMemsetAsync(dc, 0, sizeof(int32_t) * ni, s);
    (void)hipMemsetAsync(dp, 0xff, sizeof(int32_t) * ni, s);
    launch_pair_lists(s, d, h->cur, dc, dl, dp, cap);
    hipError_t e = hipMemcpyAsync(counts, dc, sizeof(int32_t) * d.n_int, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(path, dp, sizeof(int32_t) * d.n_int, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(list, dl, sizeof(int32_t) * d.n_int * (size_t)cap, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = fail(h, LSSPH_E_HIP, std::string("lssph_pair_lists: ") + hipGetErrorString(e));
  }
  if (dc) (void)hipFree(dc);
  if (dl) (void)hipFree(dl);
  if (dp) (void)hipFree(dp);
  if (rc) return rc;
  for (int i = 0; i < d.n_int; ++i)
    if (counts[i] > cap) return fail(h, LSSPH_E_CAPACITY, "lssph_pair_lists: a list is longer than cap");
  return 0;
}

int32_t lssph_probe(lssph_handle_t h, int32_t inst, int32_t field, int32_t q, int32_t n, const double *x,
                    const double *y, double *out) {
  if (int rc = ready(h)) return rc;
  Device &d = h->d;
  if (inst < 0 || inst >= d.n_inst) return fail(h, LSSPH_E_ARG, "lssph_probe: instance out of range");
  if (field < 0 || field > 4 || (field == 4 && !d.opt.has_tem))
    return fail(h, LSSPH_E_ARG, "lssph_probe: field must be 0..3, or 4 (tem) for target_problem 4");
  if (q != 3 && q != 6 && q != 10) return fail(h, LSSPH_E_ARG, "lssph_probe: q must be 3, 6 or 10");
  if (n < 0 || (n > 0 && (!x || !y || !out))) return fail(h, LSSPH_E_ARG, "lssph_probe: null array or n < 0");
  if (n == 0) return 0;
  hipStream_t s = h->stream;
  double *buf = nullptr;
  if (hipMalloc(&buf, sizeof(double) * 5 * (size_t)n) != hipSuccess)
    return fail(h, LSSPH_E_NOMEM, "lssph_probe: hipMalloc failed");
  hipError_t e = hipMemcpyAsync(buf, x, sizeof(double) * n, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(buf + n, y, sizeof(double) * n, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    launch_probe(s, d, h->cur, inst, field, q, n, buf, buf + n, buf + 2 * n);
    e = hipMemcpyAsync(out, buf + 2 * n, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(h, LSSPH_E_HIP, std::string("lssph_probe: ") + hipGetErrorString(e));
  return check_counters(h);
}

int32_t lssph_probe_static(lssph_handle_t h, int32_t inst, int32_t field, int32_t q, int32_t n, const double *x,
                           const double *y, double *out) {
  if (int rc = ready(h)) return rc;
  Device &d = h->d;
  if (inst < 0 || inst >= d.n_inst) return fail(h, LSSPH_E_ARG, "lssph_probe_static: instance out of range");
  if (field < 0 || field > 4 || (field == 4 && !d.opt.has_tem))
    return fail(h, LSSPH_E_ARG, "lssph_probe_static: field must be 0..3, or 4 (tem) for target_problem 4");
  if (q != 3 && q != 6 && q != 10) return fail(h, LSSPH_E_ARG, "lssph_probe_static: q must be 3, 6 or 10");
  if (n < 0 || (n > 0 && (!x || !y || !out)))
    return fail(h, LSSPH_E_ARG, "lssph_probe_static: null array or n < 0");
  if (n > kProbeStaticMax)
    return fail(h, LSSPH_E_ARG,
                "lssph_probe_static: " + std::to_string(n) + " probes exceed its staging buffer of " +
                    std::to_string(kProbeStaticMax) + " (split the call, or use lssph_probe)");
  if (n == 0) return 0;
  double *sp = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    sp = g_scratch_ptr[h->device];
  }
  if (!sp) return fail(h, LSSPH_E_STATE, "lssph_probe_static: staging buffer not set up on this device");
  hipStream_t s = h->stream;
  HIP_TRY(h, hipMemcpyAsync(sp, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HIP_TRY(h, hipMemcpyAsync(sp + kProbeStaticMax, y, sizeof(double) * n, hipMemcpyHostToDevice, s));
  launch_probe(s, d, h->cur, inst, field, q, n, sp, sp + kProbeStaticMax, sp + 2 * kProbeStaticMax);
  HIP_TRY(h, hipMemcpyAsync(out, sp + 2 * kProbeStaticMax, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
  HIP_TRY(h, hipStreamSynchronize(s));
  HIP_TRY(h, hipGetLastError());
  return check_counters(h);
}

int32_t lssph_stats(lssph_handle_t h, lssph_stats_t *out) {
  if (!h || !out) return fail(h, LSSPH_E_ARG, "lssph_stats: null argument");
  int32_t c[kNumCounters];
  unsigned long long s64[kNumStats64];
  HIP_TRY(h, hipMemcpyAsync(c, h->d.counters, sizeof c, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipMemcpyAsync(s64, h->d.stats64, sizeof s64, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  out->fp64_fallbacks = (int64_t)s64[kFallbacks];
  out->list_overflows = (int64_t)s64[kOverflows];
  out->max_neighbors = c[kMaxNbr];
  out->max_cell_count = c[kMaxCell];
  out->escaped = c[kEscaped];
  out->pad = 0;
  return 0;
}

int32_t lssph_reset_stats(lssph_handle_t h) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  launch_fill_u64(h->stream, h->d.stats64, 0ull, kNumStats64);
  launch_fill_i32(h->stream, h->d.counters + kMaxCell, 0, kMaxNbr - kMaxCell + 1);
  HIP_TRY(h, hipGetLastError());
  return 0;
}

int32_t lssph_synchronize(lssph_handle_t h) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  HIP_TRY(h, hipGetLastError());
  return check_counters(h);
}

int32_t lssph_set_tuning(lssph_handle_t h, int32_t block, int32_t list_len, int32_t reorder_interval) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  // every check before any change: a rejected call leaves the handle as it was
  if (block && (block < 64 || block > 256 || block % 64))
    return fail(h, LSSPH_E_ARG, "block must be 64, 128, 192 or 256");
  if (list_len && (list_len < 8 || list_len > 128)) return fail(h, LSSPH_E_ARG, "list_len must be in [8, 128]");
  const int nb = block ? block : h->d.block, nl = list_len ? list_len : h->d.list_len;
  if ((size_t)(nl + 1) * nb * 4 > 64 * 1024)
    return fail(h, LSSPH_E_ARG, "list_len * block exceeds 64 KiB of LDS (the list length kept when list_len is 0)");
  h->d.block = nb;
  h->d.list_len = nl;
  if (reorder_interval) h->reorder_interval = reorder_interval < 0 ? 0 : reorder_interval;
  drop_graphs(h);
  return 0;
}

int32_t lssph_set_skin(lssph_handle_t h, double skin) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  if (!(skin >= 0.0) || skin > 1.0) return fail(h, LSSPH_E_ARG, "lssph_set_skin: skin must be in [0, 1]");
  if (h->uploaded) return fail(h, LSSPH_E_STATE, "lssph_set_skin: call before lssph_upload");
  const double old = h->skin;
  h->skin = skin;
  if (setup_grid(h) > (1ll << 30)) {  // narrower cells: the grid lssph_create checked can grow
    h->skin = old;
    setup_grid(h);
    return fail(h, LSSPH_E_ARG, "lssph_set_skin: problem too large for 32-bit indices at this skin");
  }
  drop_graphs(h);
  if (int rc = alloc_grid(h)) return rc;
  HIP_TRY(h, hipGetLastError());
  return 0;
}

int32_t lssph_set_cell_split(lssph_handle_t h, int32_t columns) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  if (columns < 1 || columns > 4) return fail(h, LSSPH_E_ARG, "lssph_set_cell_split: columns must be 1 to 4");
  if (h->uploaded) return fail(h, LSSPH_E_STATE, "lssph_set_cell_split: call before lssph_upload");
  const int old = h->cell_split;
  h->cell_split = columns;
  if (setup_grid(h) > (1ll << 30)) {
    h->cell_split = old;
    setup_grid(h);
    return fail(h, LSSPH_E_ARG, "lssph_set_cell_split: problem too large for 32-bit indices with these cells");
  }
  drop_graphs(h);
  if (int rc = alloc_grid(h)) return rc;
  HIP_TRY(h, hipGetLastError());
  return 0;
}

int32_t lssph_set_isolation(lssph_handle_t h, int32_t on) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  if (on != 0 && on != 1) return fail(h, LSSPH_E_ARG, "lssph_set_isolation: on must be 0 or 1");
  Device &d = h->d;
  if ((d.retired != nullptr) == (on != 0)) return 0;
  if (on) {
    // a fresh start: no instance retired, steps counted from here
    HIP_TRY(h, hipMemcpyAsync(d.inst, h->inst_host.data(), sizeof(InstDev) * d.n_inst, hipMemcpyHostToDevice,
                              h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    launch_fill_i32(h->stream, d.retired_buf, 0, 2 * d.n_inst);
    launch_fill_i32(h->stream, d.counters + kStepNo, 0, 1);
    HIP_TRY(h, hipGetLastError());
  }
  d.retired = on ? d.retired_buf : nullptr;
  drop_graphs(h);
  return 0;
}

int32_t lssph_instance_status(lssph_handle_t h, int32_t *retired, int32_t *reason) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  if (!retired) return fail(h, LSSPH_E_ARG, "lssph_instance_status: null array");
  Device &d = h->d;
  HIP_TRY(h, hipMemcpyAsync(retired, d.retired_buf, sizeof(int32_t) * d.n_inst, hipMemcpyDeviceToHost, h->stream));
  if (reason)
    HIP_TRY(h, hipMemcpyAsync(reason, d.retired_buf + d.n_inst, sizeof(int32_t) * d.n_inst, hipMemcpyDeviceToHost,
                              h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return 0;
}

int32_t lssph_set_wave_particles(lssph_handle_t h, int32_t ppw) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  if (!(ppw >= 0 && ppw <= 3)) return fail(h, LSSPH_E_ARG, "lssph_set_wave_particles: 0, 1, 2 or 3");
  h->d.wave_ppw = ppw;
  drop_graphs(h);
  return 0;
}

int32_t lssph_set_scan(lssph_handle_t h, int32_t staged) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  if (staged != 0 && staged != 1) return fail(h, LSSPH_E_ARG, "lssph_set_scan: staged must be 0 or 1");
  h->d.staged_scan = staged;
  drop_graphs(h);
  return 0;
}

void *lssph_stream(lssph_handle_t h) { return h ? (void *)h->stream : nullptr; }

int32_t lssph_profile(lssph_handle_t h, int32_t enable) {
  if (!h) return fail(nullptr, LSSPH_E_ARG, "null handle");
  h->profiling = enable != 0;
  return 0;
}

int32_t lssph_kernel_times(lssph_handle_t h, lssph_kernel_times_t *out) {
  if (!h || !out) return fail(h, LSSPH_E_ARG, "lssph_kernel_times: null argument");
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  for (auto &e : h->ev_pending) {
    float ms = 0.0f;
    HIP_TRY(h, hipEventElapsedTime(&ms, e.a, e.b));
    h->kt_ms[e.group] += ms;
    h->kt_n[e.group] += 1;
    h->ev_pool.push_back(e.a);
    h->ev_pool.push_back(e.b);
  }
  h->ev_pending.clear();
  out->lsa_ms = h->kt_ms[kGrpLsa];
  out->pst_ms = h->kt_ms[kGrpPst];
  out->cells_ms = h->kt_ms[kGrpCells];
  out->ghost_ms = h->kt_ms[kGrpGhost];
  out->other_ms = h->kt_ms[kGrpOther];
  out->lsa_launches = h->kt_n[kGrpLsa];
  out->pst_launches = h->kt_n[kGrpPst];
  out->cells_launches = h->kt_n[kGrpCells];
  out->ghost_launches = h->kt_n[kGrpGhost];
  out->other_launches = h->kt_n[kGrpOther];
  for (int g = 0; g < 5; ++g) {
    h->kt_ms[g] = 0.0;
    h->kt_n[g] = 0;
  }
  return 0;
}

}  // extern "C"
