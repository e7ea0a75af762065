// pair_common.hpp -- neighbor traversal shared by the pair kernels.
//
// Neighbor data comes from the records in slot order (cells consecutive,
// each cell's particles in ascending upload index like the reference's
// lists, update_cell_mod.f90:57-92), so the stencil of a particle (three
// cell rows of 2 kx + 1 columns) is at most three contiguous slot ranges
// (four to six when a periodic row wraps) and lanes of a wavefront read the
// same or adjacent cache lines.  Cells are h_eff + skin tall and (h_eff +
// skin) / kx wide and rebuilt once per time step; every pass still tests
// r_ij^2 <= h_eff^2 at the current positions, so the neighbor sets are the
// reference's.
#pragma once

#include "internal.hpp"

namespace lssph {

// A lane's neighbor list in the block's LDS.  Each wavefront owns
// list_len + 1 rows of 256 bytes (4 * (list_len + 1) * blockDim bytes per
// block in all), and lane l's entry k sits at byte 256 k + 4 l of its
// wave's rows (consecutive lanes, consecutive banks).  Row list_len is the
// scan's trash entry.  A wave whose scan was staged (scan_staged) stores
// 16-bit entries instead, slot - base, at byte 128 k + 2 l, and uses the
// upper half of its rows to stage candidate positions.
struct NbList {
  char *wbase;   // the wave's rows
  int lane;      // lane within the wavefront
  int list_len;  // entries before the trash entry
  int cnt;       // entries, the particle itself included; > list_len: overflowed
  int base;      // >= 0: 16-bit entries relative to base; -1: 32-bit entries
  bool pos;      // 16-bit entries are row << 14 | staging index: positions stay in LDS
  unsigned long long rowd;  // with pos: slot = base + index + 16-bit field `row` of rowd
};

__device__ __forceinline__ NbList make_list(int32_t *lds, int list_len, int tid) {
  NbList L;
  L.wbase = reinterpret_cast<char *>(lds) + (size_t)(tid >> 6) * 256u * (unsigned)(list_len + 1);
  L.lane = tid & 63;
  L.list_len = list_len;
  L.cnt = 0;
  L.base = -1;
  L.pos = false;
  L.rowd = 0;
  return L;
}

// A list that is never read: every walk takes the direct stencil walk.
__device__ __forceinline__ NbList no_list() {
  NbList L;
  L.wbase = nullptr;
  L.lane = 0;
  L.list_len = 0;
  L.cnt = 1 << 30;
  L.base = -1;
  L.pos = false;
  L.rowd = 0;
  return L;
}

// Minimum and maximum of a value over a full wavefront (every lane active),
// by DPP: row shifts by 1, 2, 4 and 8 leave each 16-lane row's extreme in its
// lane 15, and the row broadcasts 15 and 31 carry those on to lane 63, whose
// value is returned (wave-uniform).  Lanes without a source in a step keep
// the identity.
__device__ __forceinline__ int wave_min_full(int v) {
  constexpr int kId = 0x7fffffff;
  v = min(v, __builtin_amdgcn_update_dpp(kId, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = min(v, __builtin_amdgcn_update_dpp(kId, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = min(v, __builtin_amdgcn_update_dpp(kId, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = min(v, __builtin_amdgcn_update_dpp(kId, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = min(v, __builtin_amdgcn_update_dpp(kId, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = min(v, __builtin_amdgcn_update_dpp(kId, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ int wave_max_full(int v) {
  constexpr int kId = -0x7fffffff - 1;
  v = max(v, __builtin_amdgcn_update_dpp(kId, v, 0x111, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(kId, v, 0x112, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(kId, v, 0x114, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(kId, v, 0x118, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(kId, v, 0x142, 0xa, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(kId, v, 0x143, 0xc, 0xf, false));
  return __builtin_amdgcn_readlane(v, 63);
}

// Orders the LDS accesses of the lanes of one wavefront (LDS operations of
// a wavefront complete in issue order; the fences keep the compiler from
// moving accesses across this point).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Image shift of a wrapped periodic cell: code 1 -> -len, 2 -> +len.
__device__ __forceinline__ double shift_of(int code, double len) {
  return code == 1 ? -len : (code == 2 ? len : 0.0);
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

// Visits the slot ranges of the stencil around local cell `lc` (rows dy =
// -1..1, within a row the columns dx = -kx..kx, the reference's order when
// kx = 1), passing the periodic image code of each range (x code in bits
// 0-1, y code in 2-3) and its range id 2 * (dy + 1) + segment.  A row is one
// range, or two when a periodic row wraps; the loops are kept rolled so the
// candidate loop body exists once in the code.
template <typename F>
__device__ __forceinline__ void visit_cells(const InstDev &I, int lc, const int32_t *__restrict__ cstart,
                                            F &&f) {
  const int nx = I.nx, ny = I.ny, kx = I.kx;
  const int cx = lc % nx, cy = lc / nx;
#pragma unroll 1
  for (int dy = -1; dy <= 1; ++dy) {
    int yy = cy + dy, ycode = 0;
    if (I.periodic) {
      if (yy < 0) {
        yy += ny;
        ycode = 1;
      } else if (yy >= ny) {
        yy -= ny;
        ycode = 2;
      }
    } else if (yy < 0 || yy >= ny) {
      continue;
    }
    const int row = I.cell_off + yy * nx;
    int a0, a1, b0 = 0, b1 = 0, ca = 0, cb = 0;
    if (!I.periodic) {
      const int x0 = max(cx - kx, 0), x1 = min(cx + kx, nx - 1);
      a0 = cstart[row + x0];
      a1 = cstart[row + x1 + 1];
    } else if (cx < kx) {
      // columns nx + cx - kx .. nx - 1 seen across the left edge, then 0 .. cx + kx
      a0 = cstart[row + nx + cx - kx];
      a1 = cstart[row + nx];
      ca = 1;
      b0 = cstart[row];
      b1 = cstart[row + cx + kx + 1];
    } else if (cx > nx - 1 - kx) {
      // columns cx - kx .. nx - 1, then 0 .. cx + kx - nx across the right edge
      a0 = cstart[row + cx - kx];
      a1 = cstart[row + nx];
      b0 = cstart[row];
      b1 = cstart[row + cx + kx + 1 - nx];
      cb = 2;
    } else {
      a0 = cstart[row + cx - kx];
      a1 = cstart[row + cx + kx + 1];
    }
#pragma unroll 1
    for (int seg = 0; seg < 2; ++seg) {
      const int t0 = seg ? b0 : a0, t1 = seg ? b1 : a1;
      f(t0, t1, (ycode << 2) | (seg ? cb : ca), 2 * (dy + 1) + seg);
    }
  }
}

// True when the lane's stencil wraps across a periodic boundary, i.e. some
// candidate range carries an image shift.
__device__ __forceinline__ bool stencil_wraps(const InstDev &I, int lc) {
  if (!I.periodic) return false;
  const int cx = lc % I.nx, cy = lc / I.nx;
  return cx < I.kx || cx > I.nx - 1 - I.kx || cy == 0 || cy == I.ny - 1;
}

// True when any active lane of the wavefront has a wrapping stencil.  Waves
// with none take loops without the image-shift arithmetic; x_j + 0 == x_j,
// so both loops form bit-identical offsets.
__device__ __forceinline__ bool wave_wraps(const InstDev &I, int lc) {
  return __ballot(stencil_wraps(I, lc)) != 0ull;
}

// Periodic image of x_j seen from x_i: the shift (-L, 0 or +L) that brings
// x_j within L/2 of x_i.  Between cell builds particles may have wrapped
// around the box, so the shift is taken per pair from the positions, and
// offsets are formed as (x_j + shift) - x_i (the oracle's expression; for
// pairs within h_eff < L/3 its cell-stencil shift is the same).
__device__ __forceinline__ double image_shift(double xj, double xi, double len) {
  const double d = xj - xi;
  return d > 0.5 * len ? -len : (d < -0.5 * len ? len : 0.0);
}

__device__ __forceinline__ int image_code(double sx, double sy) {
  return (sx < 0.0 ? 1 : (sx > 0.0 ? 2 : 0)) | ((sy < 0.0 ? 1 : (sy > 0.0 ? 2 : 0)) << 2);
}

// Candidate scan of the particle at (xi, yi) within radius^2 r2: fills the
// LDS list (slot | image code << kSlotBits) in stencil order and returns the
// count.  The particle's own slot is among the candidates and is accepted
// (r = 0): it takes one list entry, and a pair operator sees it as a pair
// with zero offset (the LS basis vanishes there, so it adds exactly zero to
// the sums).  The loop body has no branch: every candidate is written at the
// current list position (position list_len is a trash entry) and the
// position advances when the candidate is accepted.  Waves in which no
// lane's stencil wraps skip the image arithmetic (x_j + 0 == x_j, so both
// paths form bit-identical offsets).  Four candidates per iteration (two in
// the wrapping waves), so several record loads are in flight before the
// first test.
__device__ __forceinline__ int scan_candidates(const Rec *__restrict__ mir, const int32_t *__restrict__ cstart,
                                               const InstDev &I, int lc, double xi, double yi, double r2,
                                               const NbList &L) {
  int cnt = 0;
  char *lbase = L.wbase + 4 * L.lane;
  const unsigned list_len = (unsigned)L.list_len;
  auto accept = [&](int t, double dx, double dy, int code) {
    const unsigned k = min((unsigned)cnt, list_len);
    *reinterpret_cast<int32_t *>(lbase + (k << 8)) = t | (code << kSlotBits);
    cnt += (sq_dist(dx, dy) <= r2) ? 1 : 0;
  };
  if (!wave_wraps(I, lc)) {
    visit_cells(I, lc, cstart, [&](int t0, int t1, int, int) {
      int t = t0;
      for (; t + 4 <= t1; t += 4) {
        const double2 p0 = load_xy(mir, t), p1 = load_xy(mir, t + 1);
        const double2 p2 = load_xy(mir, t + 2), p3 = load_xy(mir, t + 3);
        accept(t, p0.x - xi, p0.y - yi, 0);
        accept(t + 1, p1.x - xi, p1.y - yi, 0);
        accept(t + 2, p2.x - xi, p2.y - yi, 0);
        accept(t + 3, p3.x - xi, p3.y - yi, 0);
      }
      for (; t < t1; ++t) {
        const double2 p0 = load_xy(mir, t);
        accept(t, p0.x - xi, p0.y - yi, 0);
      }
    });
  } else {
    auto test = [&](int t, double2 p) {
      const double sx = image_shift(p.x, xi, I.len_x), sy = image_shift(p.y, yi, I.len_y);
      accept(t, (p.x + sx) - xi, (p.y + sy) - yi, image_code(sx, sy));
    };
    visit_cells(I, lc, cstart, [&](int t0, int t1, int, int) {
      int t = t0;
      for (; t + 2 <= t1; t += 2) {
        const double2 p0 = load_xy(mir, t), p1 = load_xy(mir, t + 1);
        test(t, p0);
        test(t + 1, p1);
      }
      if (t < t1) test(t, load_xy(mir, t));
    });
  }
  return cnt;
}

// Wave-staged candidate scan.  The lanes of a full wavefront whose
// stencils share one instance and at most two adjacent cell rows and do not
// wrap read the union of their candidate rows (three or four contiguous
// slot ranges) once, coalesced, into LDS, and each lane then tests its own
// window from there: one 16-byte load per lane per 64 candidates instead of
// one per lane per candidate.  When all rows fit the staging area at once,
// the positions stay there for the walk (list entries row << 14 | staging
// index, L.pos); otherwise the rows are staged one after another (entries
// slot - base).  Either way the accepted entries and their order are those
// of scan_candidates.  Returns false, with the list untouched, when the wave
// does not qualify (partial or mixed waves, periodic edges, a row union
// larger than the staging area, a slot span beyond 16 bits).
__device__ __forceinline__ bool scan_staged(const Rec *__restrict__ mir, const int32_t *__restrict__ cstart,
                                            const InstDev &I, int lc, double xi, double yi, double r2,
                                            NbList &L) {
  if (__builtin_amdgcn_read_exec() != ~0ull) return false;
  const int nx = __builtin_amdgcn_readfirstlane(I.nx);
  const int ny = __builtin_amdgcn_readfirstlane(I.ny);
  const int kx = __builtin_amdgcn_readfirstlane(I.kx);
  const int off = __builtin_amdgcn_readfirstlane(I.cell_off);
  const int cx = lc % I.nx, cy = lc / I.nx;
  if (__ballot(I.cell_off != off || stencil_wraps(I, lc)) != 0ull) return false;
  // particles that crossed into the next cell row since the last build put
  // a wave on two rows: four candidate rows, three of them per lane
  const int cy0 = wave_min_full(cy), cy1 = wave_max_full(cy);
  if (cy1 - cy0 > 1) return false;
  const int x0 = max(wave_min_full(cx) - kx, 0);
  const int x1 = min(wave_max_full(cx) + kx, nx - 1);
  const int y0 = max(cy0 - 1, 0), y1 = min(cy1 + 1, ny - 1);
  const int cap = 8 * (L.list_len + 1);  // 16-byte positions in the upper half of the wave's rows
  const int base = cstart[off + y0 * nx + x0];
  if (cstart[off + y1 * nx + x1 + 1] - base > 0xffff) return false;
  int total = 0;
  for (int yy = y0; yy <= y1; ++yy) {
    const int n_row = cstart[off + yy * nx + x1 + 1] - cstart[off + yy * nx + x0];
    if (n_row > cap) return false;
    total += n_row;
  }

  const int lane = L.lane;
  char *l16 = L.wbase + 2 * lane;
  double2 *stage = reinterpret_cast<double2 *>(L.wbase + 128 * (L.list_len + 1));
  const unsigned list_len = (unsigned)L.list_len;
  const int lx0 = max(cx - kx, 0), lx1 = min(cx + kx, nx - 1);
  int cnt = 0;
  auto scan_window = [&](int a0, int a1, int shift, auto &&entry) {
    auto accept = [&](int t, double2 p) {
      const unsigned k = min((unsigned)cnt, list_len);
      *reinterpret_cast<uint16_t *>(l16 + (k << 7)) = entry(t);
      cnt += (sq_dist(p.x - xi, p.y - yi) <= r2) ? 1 : 0;
    };
    int t = a0;
    for (; t + 4 <= a1; t += 4) {
      const double2 *sp = stage + (t - shift);
      const double2 p0 = sp[0], p1 = sp[1], p2 = sp[2], p3 = sp[3];
      accept(t, p0);
      accept(t + 1, p1);
      accept(t + 2, p2);
      accept(t + 3, p3);
    }
    for (; t < a1; ++t) accept(t, stage[t - shift]);
  };
  if (total <= cap) {
    // every row at once: the positions stay in LDS for the walk, and a list
    // entry is row << 14 | staging index (rows 0..3, index < cap <= 1032)
    unsigned long long rowd = 0;
    int seg = 0;
    for (int yy = y0, r = 0; yy <= y1; ++yy, ++r) {
      const int row = off + yy * nx;
      const int u0 = cstart[row + x0], n_row = cstart[row + x1 + 1] - u0;
      for (int i = lane; i < n_row; i += 64) stage[seg + i] = load_xy(mir, u0 + i);
      // rows are disjoint ascending slot ranges: 0 <= (u0 - base) - seg <= 0xffff
      rowd |= (unsigned long long)((u0 - base) - seg) << (16 * r);
      seg += n_row;
    }
    wave_sync();
    seg = 0;
#pragma unroll 1
    for (int yy = y0, r = 0; yy <= y1; ++yy, ++r) {
      const int row = off + yy * nx;
      const int u0 = cstart[row + x0], n_row = cstart[row + x1 + 1] - u0;
      const int a0 = cstart[row + lx0];
      const int a1 = (yy >= cy - 1 && yy <= cy + 1) ? cstart[row + lx1 + 1] : a0;
      const int shift = u0 - seg, tag = (r << 14) - shift;
      scan_window(a0, a1, shift, [&](int t) { return (uint16_t)(t + tag); });
      seg += n_row;
    }
    L.pos = true;
    L.rowd = rowd;
  } else {
    // one row at a time; a list entry is slot - base
#pragma unroll 1
    for (int yy = y0; yy <= y1; ++yy) {
      const int row = off + yy * nx;
      const int u0 = cstart[row + x0], u1 = cstart[row + x1 + 1];
      const int a0 = cstart[row + lx0];
      const int a1 = (yy >= cy - 1 && yy <= cy + 1) ? cstart[row + lx1 + 1] : a0;
      wave_sync();  // the previous row's reads of the staging area come first
      for (int i = lane; i < u1 - u0; i += 64) stage[i] = load_xy(mir, u0 + i);
      wave_sync();
      scan_window(a0, a1, u0, [&](int t) { return (uint16_t)(t - base); });
    }
    L.pos = false;
  }
  L.cnt = cnt;
  L.base = base;
  return true;
}

// Fills the lane's list: the staged scan when it is allowed and the wave
// qualifies, else the per-lane scan.
__device__ __forceinline__ void build_list(const Rec *__restrict__ mir, const int32_t *__restrict__ cstart,
                                           const InstDev &I, int lc, double xi, double yi, double r2, NbList &L,
                                           bool staged) {
  if (!staged || !scan_staged(mir, cstart, I, lc, xi, yi, r2, L)) {
    L.base = -1;
    L.cnt = scan_candidates(mir, cstart, I, lc, xi, yi, r2, L);
  }
}

// Calls f(nb_record, dx, dy, slot) for every neighbor within h_eff of slot
// s, in stencil order, by a walk over the candidates of the cell list.
template <typename F>
__device__ __forceinline__ void walk_stencil(const Rec *__restrict__ mir, const int32_t *__restrict__ cstart,
                                             const InstDev &I, int lc, int s, const Rec &me, F &&f) {
  const bool per = I.periodic;
  visit_cells(I, lc, cstart, [&](int t0, int t1, int, int) {
    for (int t = t0; t < t1; ++t) {
      if (t == s) continue;
      const double2 xy = load_xy(mir, t);
      const double sx = per ? image_shift(xy.x, me.x, I.len_x) : 0.0;
      const double sy = per ? image_shift(xy.y, me.y, I.len_y) : 0.0;
      const double dx = (xy.x + sx) - me.x, dy = (xy.y + sy) - me.y;
      if (sq_dist(dx, dy) <= I.h_eff_sq) f(mir[t], dx, dy, t);
    }
  });
}

// Loop over k = 0 .. n-1 calling body(k); with DEEP the loop is unrolled
// four times so that four neighbors' loads are in flight before their sums
// (still added in list order), else the compiler picks the unrolling.  The
// walk waits on gathers more than on arithmetic, but the deeper loop needs
// registers that the heavier kernels do not have.
template <bool DEEP, typename B>
__device__ __forceinline__ void list_loop(int n, B &&body) {
  if constexpr (DEEP) {
#pragma unroll 4
    for (int k = 0; k < n; ++k) body(k);
  } else {
    for (int k = 0; k < n; ++k) body(k);
  }
}

// Calls f(nb_record, dx, dy, slot) for every accepted neighbor of slot s in
// summation order: from the LDS list filled by scan_candidates or
// scan_staged, which holds the particle itself as a pair with dx = dy = 0,
// or (the list overflowed) by a walk over the candidates, which skips it.
template <bool DEEP = false, typename F>
__device__ __forceinline__ void walk_neighbors(const Rec *__restrict__ mir, const int32_t *__restrict__ cstart,
                                               const InstDev &I, int lc, int s, const Rec &me, const NbList &L,
                                               F &&f) {
  const int cnt = L.cnt;
  if (cnt > L.list_len) {
    walk_stencil(mir, cstart, I, lc, s, me, f);
    return;
  }
  if (L.base >= 0 && L.pos) {
    // positions from the staging area, the other fields from the records
    const char *l16 = L.wbase + 2 * L.lane;
    const double2 *stage = reinterpret_cast<const double2 *>(L.wbase + 128 * (L.list_len + 1));
    list_loop<DEEP>(cnt, [&](int k) {
      const unsigned e = *reinterpret_cast<const uint16_t *>(l16 + (k << 7));
      const unsigned i = e & 0x3fffu;
      // (e >> 14) * 16 = (e >> 10) & 0x30, as i < 1024 + 8
      const unsigned d = (unsigned)(L.rowd >> ((e >> 10) & 0x30u)) & 0xffffu;
      const double2 p = stage[i];
      const int t = L.base + (int)(i + d);
      const Rec nb = mir[t];
      f(nb, p.x - me.x, p.y - me.y, t);
    });
  } else if (L.base >= 0) {
    const char *l16 = L.wbase + 2 * L.lane;
    list_loop<DEEP>(cnt, [&](int k) {
      const int t = L.base + *reinterpret_cast<const uint16_t *>(l16 + (k << 7));
      const Rec nb = mir[t];
      f(nb, nb.x - me.x, nb.y - me.y, t);
    });
  } else if (!wave_wraps(I, lc)) {
    const char *l32 = L.wbase + 4 * L.lane;
    for (int k = 0; k < cnt; ++k) {
      const int t = *reinterpret_cast<const int32_t *>(l32 + (k << 8)) & kSlotMask;
      const Rec nb = mir[t];
      f(nb, nb.x - me.x, nb.y - me.y, t);
    }
  } else {
    const char *l32 = L.wbase + 4 * L.lane;
    for (int k = 0; k < cnt; ++k) {
      const int e = *reinterpret_cast<const int32_t *>(l32 + (k << 8));
      const int t = e & kSlotMask, code = e >> kSlotBits;
      const Rec nb = mir[t];
      const double dx = (nb.x + shift_of(code & 3, I.len_x)) - me.x;
      const double dy = (nb.y + shift_of(code >> 2, I.len_y)) - me.y;
      f(nb, dx, dy, t);
    }
  }
}

// Cell coordinates on the internal grid: cells of height h_eff + skin and
// width (h_eff + skin) / kx (floor(x / width)), so that the stencil (rows
// -1..1, columns -kx..kx) of a cell list built at one position still holds
// every pair within h_eff after each particle has moved by up to skin / 2.  Periodic boxes clamp the upper edge.  Returns false
// outside the grid, also for a non-finite position (the range test is made
// on the floating-point cell coordinate, before any conversion).
__device__ __forceinline__ bool fluid_cell_xy(const InstDev &I, double x, double y, int &cx, int &cy) {
  double fx = floor(x * I.inv_cw_x), fy = floor(y * I.inv_cw_y);
  if (I.periodic) {
    fx = (fx == (double)I.nx) ? (double)(I.nx - 1) : fx;
    fy = (fy == (double)I.ny) ? (double)(I.ny - 1) : fy;
  }
  const bool in = fx >= 0.0 && fx < (double)I.nx && fy >= 0.0 && fy < (double)I.ny;
  cx = in ? (int)fx : -1;
  cy = in ? (int)fy : -1;
  return in;
}

// Local cell index on the internal grid, -1 outside.
__device__ __forceinline__ int fluid_cell(const InstDev &I, double x, double y) {
  int cx, cy;
  if (!fluid_cell_xy(I, x, y, cx, cy)) return -1;
  return cx + cy * I.nx;
}

// Same, clamped into the grid (used after binning has flagged escapes).
__device__ __forceinline__ int local_cell(const InstDev &I, double x, double y) {
  const int c = fluid_cell(I, x, y);
  return c < 0 ? 0 : c;
}

// Walls and virtual markers use the same internal grid.
__device__ __forceinline__ int fixed_cell(const InstDev &I, double x, double y) { return fluid_cell(I, x, y); }

// The reference's 1-based cell index (cell%idx_SP): fluid particles
// ceiling(x * inv_h_eff) (update_cell_mod.f90:53-55), walls and markers
// ceiling(x / h_eff) (init_cell_mod.f90:76-78); periodic boxes (extension)
// floor(x * cells / L).
__device__ __forceinline__ int ref_cell_index(const InstDev &I, double x, double y, bool wall) {
  int cx, cy;
  if (I.periodic) {
    cx = (int)floor(x * I.inv_rcw_x);
    cy = (int)floor(y * I.inv_rcw_y);
    cx = (cx == I.rnx) ? I.rnx - 1 : cx;
    cy = (cy == I.rny) ? I.rny - 1 : cy;
    return cx + 1 + cy * I.rnx;
  }
  if (wall) {
    cx = (int)ceil(x / I.h_eff);
    cy = (int)ceil(y / I.h_eff);
  } else {
    cx = (int)ceil(x * I.inv_h_eff);
    cy = (int)ceil(y * I.inv_h_eff);
  }
  return cx + (cy - 1) * I.rnx;
}

// Reference cell coordinates (0-based) of a position: see ref_cell_index.
__device__ __forceinline__ void ref_cell_xy(const InstDev &I, double x, double y, bool wall, int &cx, int &cy) {
  if (I.periodic) {
    cx = (int)floor(x * I.inv_rcw_x);
    cy = (int)floor(y * I.inv_rcw_y);
    cx = (cx == I.rnx) ? I.rnx - 1 : cx;
    cy = (cy == I.rny) ? I.rny - 1 : cy;
  } else if (wall) {
    cx = (int)ceil(x / I.h_eff) - 1;
    cy = (int)ceil(y / I.h_eff) - 1;
  } else {
    cx = (int)ceil(x * I.inv_h_eff) - 1;
    cy = (int)ceil(y * I.inv_h_eff) - 1;
  }
}

// Bookkeeping of a moved particle at slot s between cell builds.
//   guard:  the particle may move at most skin / 2 from where the current
//           cell list was built (xbuild); beyond that the stencil could
//           miss pairs, so the move is counted as an error;
//   keys:   in the pass right before a build, the new cell key goes to
//           kslot[s] (which held the key at the build); an escape is counted
//           and clamped into the grid like the full build does; a move that
//           leaves the stencil of the old cell (more than kx columns or one
//           row) is queued on the stray list.
// A non-finite position fails both checks.  In isolation mode (retired
// non-null) such a move retires the particle's instance instead of raising
// an error: the particle keeps its cell key, and from the next pass on the
// kernels leave the instance's records as they are.  Otherwise the move is
// counted as the handle's error, and the particle and its instance are
// frozen the same way until the next upload.  A diverging run would
// otherwise step on until the next synchronising call reports the error,
// binning every escaped particle into one cell: that cell would grow by up
// to a stray list per build, and its members' scans and its rebuild would
// grow with the square of its size.
struct MoveOut {
  int32_t *kslot;
  int32_t *strays;
  int32_t *nstray;
  int32_t *counters;
  const double *xbuild;  // [2 * n] position at the last cell build, per slot
  int guard, keys;
  int32_t *retired;      // [2 n_inst], isolation mode only: step, then reason
  const int32_t *sinfo;  // instance of a slot (read when retiring)
  int n_inst;
  InstDev *inst_rw;      // the instance constants, whose `retired` flag a retirement sets
};

// Runs run(I, ii) for the particle of info word inf with its instance's
// constants, or keep() when the instance has been retired (isolation mode).
// A wave of one instance (the common case) reads the constants with scalar
// loads; a wave that spans instances, with per-lane loads.
template <typename F, typename K>
__device__ __forceinline__ void for_instance(const InstDev *inst, int32_t inf, F &&run, K &&keep) {
  const int ii = info_inst(inf), ii0 = __builtin_amdgcn_readfirstlane(ii);
  if (__ballot(ii != ii0) == 0ull) {
    const InstDev &I = *inst_const(inst, ii0);
    if (I.retired)
      keep();
    else
      run(I, ii0);
  } else {
    const InstDev &I = inst[ii];
    if (I.retired)
      keep();
    else
      run(I, ii);
  }
}

// Marks the instance of slot s retired during the current step, with the
// cause: 1 (LSSPH_RETIRED_SKIN) a move beyond the skin, 2
// (LSSPH_RETIRED_ESCAPED) a position outside the grid or not finite.  Every
// thread that retires the instance in one pass stores the same step (later
// passes skip the instance), and the reason is one of the causes seen.
__device__ __forceinline__ void retire_instance(const MoveOut &M, int s, int reason) {
  const int ii = info_inst(M.sinfo[s]);
  M.retired[ii] = max(M.counters[kStepNo], 1);
  M.retired[M.n_inst + ii] = reason;
  M.inst_rw[ii].retired = 1;
}

__device__ __forceinline__ void record_move(const MoveOut &M, const InstDev &I, int s, double x, double y) {
  // the common case takes one branch; the rare outcomes (a move beyond the
  // skin, a position outside the grid) are handled after it
  int reason = 0;
  if (M.guard) {
    double dx = x - M.xbuild[2 * s], dy = y - M.xbuild[2 * s + 1];
    if (I.periodic) {
      dx = dx > 0.5 * I.len_x ? dx - I.len_x : (dx < -0.5 * I.len_x ? dx + I.len_x : dx);
      dy = dy > 0.5 * I.len_y ? dy - I.len_y : (dy < -0.5 * I.len_y ? dy + I.len_y : dy);
    }
    if (!(dx * dx + dy * dy <= I.skin_half_sq)) reason = LSSPH_RETIRED_SKIN;
  }
  int cx = 0, cy = 0;
  if (M.keys && !fluid_cell_xy(I, x, y, cx, cy)) reason = LSSPH_RETIRED_ESCAPED;
  if (reason != 0) {
    // the particle keeps its cell key either way
    if (M.retired) {
      retire_instance(M, s, reason);
    } else {
      if (reason == LSSPH_RETIRED_SKIN)
        M.counters[kSkinExceeded] = 1;
      else
        atomicAdd(&M.counters[kEscaped], 1);
      M.inst_rw[info_inst(M.sinfo[s])].retired = 1;  // frozen until the next upload
    }
    return;
  }
  if (!M.keys) return;
  const int nx = I.nx, ny = I.ny;
  const int oc = M.kslot[s] - I.cell_off;
  // column and row distance between the old and the new cell (across the
  // edges of a periodic box the shorter way round)
  int ddx = cx - oc % nx, ddy = cy - oc / nx;
  ddx = ddx < 0 ? -ddx : ddx;
  ddy = ddy < 0 ? -ddy : ddy;
  if (I.periodic) {
    ddx = min(ddx, nx - ddx);
    ddy = min(ddy, ny - ddy);
  }
  const int key = I.cell_off + cx + cy * nx;
  M.kslot[s] = key;
  if (ddx > I.kx || ddy > 1) {
    const int k = atomicAdd(M.nstray, 1);
    if (k < kStrayCap) {
      M.strays[2 * k] = s;
      M.strays[2 * k + 1] = key;
    } else {
      atomicAdd(&M.counters[kStrayOverflow], 1);
    }
  }
}

}  // namespace lssph
