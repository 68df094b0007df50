// HIP kernels of the mask pipeline for gfx950:
//   flatten  -- per-segment subdivision tree walk (f64, bit-exact)
//   bin      -- cell binning of the flattened lines: per-tile records, band
//               backdrops, free-endpoint events (count pass that keeps short
//               lines' records, scans, scatter)
//   fine     -- per-tile LDS accumulation of signed area/cover, DPP row scan,
//               fill rule, coalesced f32 stores
#pragma once
#include <hip/hip_runtime.h>
#include "rz_geom.hpp"
#include "rz_paint.hpp"

namespace rzk {

using rzg::Pt;

constexpr int kBlock = 256;

// Device view of a resident path. Slots are the flatten units in the
// reference's order: for every subpath its segments, then its closing line
// (rasterizers always flatten with close = true, src/rasterize.rs:307,330).
struct DPath {
  const double* pts;          // x,y pairs
  const uint8_t* kinds;       // per segment
  const uint32_t* pt_off;     // per segment
  const uint32_t* slot_item;  // slot -> segment index, or 0x80000000|subpath for a closing line
  const uint8_t* slot_flags;  // bit0: start is a gap, bit1: end is a gap
  const double* close_pts;    // per subpath: closing line (x0,y0,x1,y1) untransformed
  const uint32_t* order;      // processing order: slot indices grouped by kind, padded with ~0u
  const uint32_t* gap_slots;  // slots whose start is a gap (slot_flags bit0)
  uint32_t n_gaps;
  uint32_t n_slots;
  uint32_t blocks_line, blocks_quad, blocks_cubic;  // block ranges per kind in `order`
};

struct Status {
  uint32_t err;        // first error (rzg::Status), 0 if none
  uint32_t overflow;   // bit0 records, bit1 events, bit2 lines, bit3 scene slices,
                       // bit4 a part of the record list
  uint32_t n_records;  // total records (written by the scan)
  uint32_t n_events;
  uint32_t n_lines;    // flattened lines (atomic counter of the flatten kernel)
  uint32_t pad[3];     // a batched scene's slice totals
  uint32_t n_list;     // entries the fullest record-list part needed, when one overflowed
  uint32_t n_line_part;  // lines the fullest line part needed, when one overflowed
};
constexpr int kStatusWords = sizeof(Status) / 4;

__device__ inline void set_err(Status* st, uint32_t e) { atomicCAS(&st->err, 0u, e); }

// kind of the slot's node for the block, and the slot index handled by this thread
__device__ inline int block_kind(const DPath& p, uint32_t b) {
  if (b < p.blocks_line) return rzg::KIND_LINE;
  if (b < p.blocks_line + p.blocks_quad) return rzg::KIND_QUAD;
  return rzg::KIND_CUBIC;
}

template <int NP>
__device__ inline rzg::Node<NP> load_node(const DPath& p, uint32_t slot, const double* m) {
  rzg::Node<NP> n;
  uint32_t item = p.slot_item[slot];
  if (item & 0x80000000u) {
    const double* c = p.close_pts + 4 * (item & 0x7fffffffu);
    n.p[0] = rzg::tr_apply(m, rzg::mkpt(c[0], c[1]));
    n.p[NP - 1] = rzg::tr_apply(m, rzg::mkpt(c[2], c[3]));
  } else {
    const double* q = p.pts + 2 * (size_t)p.pt_off[item];
#pragma unroll
    for (int i = 0; i < NP; ++i) n.p[i] = rzg::tr_apply(m, rzg::mkpt(q[2 * i], q[2 * i + 1]));
  }
  return n;
}

struct Xform {
  double m[6];
};

// One path of a batched scene as the device sees it in one render: its
// transform into its sub-view (align * tr, src/scene.rs:425-429), the
// sub-view's binning grid (phased onto the canvas tile grid), the offsets of
// its slices of the per-tile, per-band and per-row arrays, and its paint
// resolved for the sub-view.
struct PathInfo {
  double m[6];
  rzg::BinGrid g;
  int live;      // has a non-empty sub-view and a drawable paint
  int rule;      // FillRule
  int ox, oy;    // sub-view origin in layer pixels
  int tx0, ty0;  // canvas tile holding the sub-view's tile (0, 0)
  uint32_t tile_base, band_base, redge_base;
  int redge_col;  // right-edge column (see k_fine)
  double redge_w;
  rzp::DPaint paint;
};

// ---------------------------------------------------------------------------
// Flatten stage hook: count leaves per slot, then emit f64 lines.
// ---------------------------------------------------------------------------
template <int NP>
__device__ void flatten_count_slot(const DPath& p, uint32_t slot, const Xform& x, double fl,
                                   uint32_t* counts, Status* st) {
  rzg::Node<NP> root = load_node<NP>(p, slot, x.m);
  uint32_t n = 0;
  int r = rzg::traverse(root, fl, [&](Pt, Pt) { ++n; });
  if (r != rzg::ST_OK) set_err(st, (uint32_t)r);
  counts[slot] = n;
}

__global__ __launch_bounds__(kBlock) void k_flatten_count(DPath p, Xform x, double fl,
                                                           uint32_t* counts, Status* st) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  uint32_t slot = p.order[i];
  if (slot == ~0u) return;
  switch (block_kind(p, blockIdx.x)) {
    case rzg::KIND_LINE: flatten_count_slot<2>(p, slot, x, fl, counts, st); break;
    case rzg::KIND_QUAD: flatten_count_slot<3>(p, slot, x, fl, counts, st); break;
    default: flatten_count_slot<4>(p, slot, x, fl, counts, st); break;
  }
}

template <int NP>
__device__ void flatten_emit_slot(const DPath& p, uint32_t slot, const Xform& x, double fl,
                                  const uint32_t* offs, double* lines) {
  rzg::Node<NP> root = load_node<NP>(p, slot, x.m);
  double* out = lines + 4 * (size_t)offs[slot];
  rzg::traverse(root, fl, [&](Pt a, Pt b) {
    out[0] = a.x;
    out[1] = a.y;
    out[2] = b.x;
    out[3] = b.y;
    out += 4;
  });
}

__global__ __launch_bounds__(kBlock) void k_flatten_emit(DPath p, Xform x, double fl,
                                                          const uint32_t* counts,
                                                          const uint32_t* offs, double* lines) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  uint32_t slot = p.order[i];
  if (slot == ~0u || counts[slot] == 0) return;
  switch (block_kind(p, blockIdx.x)) {
    case rzg::KIND_LINE: flatten_emit_slot<2>(p, slot, x, fl, offs, lines); break;
    case rzg::KIND_QUAD: flatten_emit_slot<3>(p, slot, x, fl, offs, lines); break;
    default: flatten_emit_slot<4>(p, slot, x, fl, offs, lines); break;
  }
}

// Cross-lane steps on DPP (no LDS round trip). update_dpp(old, src, ctrl,
// row_mask, bank_mask, bound_ctrl) gives `old` to lanes whose source lies
// outside their 16-lane row or whose row is masked off. Controls: row_shr:n
// 0x110 + n, wave_shl:1 0x130, wave_shr:1 0x138, row_bcast:15 0x142,
// row_bcast:31 0x143.
__device__ inline int dpp_add_scan(int x) {  // inclusive sum over the wave
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // rows 1, 3 += rows 0, 2
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // rows 2, 3 += rows 0..1
  return x;
}
__device__ inline int dpp_max_scan(int x) {  // inclusive max over the wave, x >= -1
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x111, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x112, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x114, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x118, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x142, 0xa, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x143, 0xc, 0xf, false));
  return x;
}
__device__ inline int lane_next(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false); }
__device__ inline int lane_prev(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false); }
__device__ inline double lane_next(double v) {
  return __hiloint2double(lane_next(__double2hiint(v)), lane_next(__double2loint(v)));
}
__device__ inline double lane_prev(double v) {
  return __hiloint2double(lane_prev(__double2hiint(v)), lane_prev(__double2loint(v)));
}

// ---------------------------------------------------------------------------
// Flatten to lines for binning, one kernel. Line slots take one thread each;
// every quad / cubic slot is split into 2^D0 virtual subtrees (the subtree at
// depth D0 along path bits j), one thread each: the thread walks from the
// root along j; a leaf met above depth D0 is processed only by the leftmost j
// under it. Every leaf is produced exactly once, with the same arithmetic as
// the slot-wide traversal, and a curve with hundreds of lines is spread over
// many threads.
//
// A thread keeps the first kLineKeep lines of its subtree in LDS while it
// counts them (consecutive lines share their endpoints, so kLineKeep + 1
// points); a wave then takes the range for its lines with one atomic on one
// of kLineParts counters and each thread writes its lines there in order, so
// the lines of a wave stay consecutive (binning wants neighbouring lines
// together). Only a thread with more lines walks its subtree a second time.
// For the 200k-segment path, 99.9% of the lines come from threads with at
// most 8 lines.
// ---------------------------------------------------------------------------
constexpr int kLineKeep = 8;
constexpr uint32_t kLineParts = 64;

struct LineBufs {
  double4* lines;       // [kLineParts][part_cap]
  uint32_t* line_path;  // batched scenes: path of each line
  uint32_t part_cap;    // lines per part
  uint32_t* fill;       // [kLineParts] lines per part
  uint2* wave_lines;    // per flatten wave: first line, line count
  uint32_t n_waves;     // flatten waves
  double ae_h;          // ActiveEdgeRasterizer: image rows of a single path (rzg::ae_chain)
  rzg::WrapClip wrap;   // a single path's right clip for rzg::wrap_lines
  Status* st;
};

// The right clip for rzg::wrap_lines of a line of path pid: the single path's,
// or a batched path's sub-view when it is drawn as mask_iter (scenes; the
// atlas's masks leave lines beyond their slot out).
template <bool MULTI>
__device__ inline rzg::WrapClip wrap_clip(const LineBufs& lb, const PathInfo* pi, uint32_t pid,
                                          bool active) {
  if (!MULTI) return lb.wrap;
  rzg::WrapClip q{-1.0, 1.0, 0, 0};
  if (active && pi[pid].g.mode == rzg::EDGE_ITER) {
    const rzg::BinGrid& g = pi[pid].g;
    q = rzg::WrapClip{g.wc, g.wc + 1.0, g.W, g.H};
  }
  return q;
}

// Walk from the slot's root along j for D0 levels. Returns the subtree root
// in *n and whether a traversal from it is needed; *leaf is set when a leaf
// above D0 is owned by this j.
template <int NP>
__device__ inline bool walk_to_subtree(rzg::Node<NP>* n, uint32_t j, int D0, double fl, bool* leaf,
                                       int* err) {
  *leaf = false;
  for (int k = 0; k < D0; ++k) {
    if (rzg::has_nans(*n)) {
      *err = rzg::ST_NAN;
      return false;
    }
    if (rzg::flatness(*n) < fl) {
      *leaf = (j & ((1u << (D0 - k)) - 1u)) == 0;
      return false;
    }
    rzg::Node<NP> c;
    rzg::child(*n, (int)((j >> (D0 - 1 - k)) & 1u), &c);
    *n = c;
  }
  return true;
}

// The subtree a thread of the flatten kernel handles. Returns the node to
// traverse (walk = true) or a single leaf owned by this thread (leaf = true).
template <int NP, bool MULTI>
__device__ inline void flat_subtree(const DPath& p, uint32_t slot, uint32_t j, int D0, bool active,
                                    const Xform& x, double fl, const PathInfo* pi,
                                    const uint32_t* slot_path, rzg::Node<NP>* n, bool* walk,
                                    bool* leaf, int* err, uint32_t* pid) {
  *walk = *leaf = false;
  *err = rzg::ST_OK;
  *pid = 0;
  if (MULTI && active) {
    *pid = slot_path[slot];
    active = pi[*pid].live != 0;
  }
  if (active) {
    *n = load_node<NP>(p, slot, MULTI ? pi[*pid].m : x.m);
    *walk = walk_to_subtree(n, j, D0, fl, leaf, err);
  }
}

// One flatten thread: count (keeping the first lines), take the wave's range,
// write. Every lane of the wave takes part.
// AE: ActiveEdgeRasterizer, each leaf line becomes the chain rzg::ae_chain
// makes of it (the line itself unless it comes near x = 0).
// WRAP: look for lines kept beyond the right clip (rzg::wraps); kernels for
// forms that never have them are built without it, which keeps the walk's
// registers down (the batched flatten then runs at four waves per SIMD).
template <int NP, bool MULTI, bool AE, bool WRAP>
__device__ void flat_item(const DPath& p, uint32_t slot, uint32_t j, int D0, bool active,
                          const Xform& x, double fl, const PathInfo* pi, const uint32_t* slot_path,
                          const LineBufs& lb, double2* keep, uint32_t part, uint32_t wave) {
  rzg::Node<NP> n;
  bool walk, leaf;
  int err;
  uint32_t pid;
  flat_subtree<NP, MULTI>(p, slot, j, D0, active, x, fl, pi, slot_path, &n, &walk, &leaf, &err, &pid);
  double ae_h = 0.0;
  if (AE) ae_h = MULTI ? (double)(active ? pi[pid].g.H : 0) : lb.ae_h;
  const rzg::WrapClip wq =
      WRAP ? wrap_clip<MULTI>(lb, pi, pid, active) : rzg::WrapClip{-1.0, 1.0, 0, 0};
  // each line of the subtree, as it is, as its chain or as the deposits of a
  // line kept beyond the right clip (then the lines are not a chain, and the
  // kept points below cannot stand for them)
  bool chain = true;
  int werr = rzg::ST_OK;
  auto lines_of = [&](Pt a, Pt c, auto&& f) {
    if (AE) {
      rzg::ae_chain(a, c, ae_h, f);
    } else if (rzg::wraps(a, c, wq)) {
      chain = false;
      if (rzg::wrap_lines(a, c, wq, [&](Pt u, Pt v) { f(u, rzg::mkpt(rzg::wrap_tag(), v.y)); }))
        werr = rzg::ST_WRAP;
    } else {
      f(a, c);
    }
  };
  uint32_t cnt = 0;
  auto count = [&](Pt a, Pt c) {
    if (cnt < (uint32_t)kLineKeep) {
      if (cnt == 0) keep[0] = make_double2(a.x, a.y);
      keep[cnt + 1] = make_double2(c.x, c.y);
    }
    ++cnt;
  };
  if (leaf) lines_of(n.p[0], n.p[NP - 1], count);
  if (walk) {
    const int r = rzg::traverse(n, fl, [&](Pt a, Pt c) { lines_of(a, c, count); });
    if (r != rzg::ST_OK) err = r;
  }
  if (err == rzg::ST_OK) err = werr;
  if (err != rzg::ST_OK) {
    set_err(lb.st, (uint32_t)err);
    cnt = 0;
  }
  const int lane = threadIdx.x & 63;
  const uint32_t incl = (uint32_t)dpp_add_scan((int)cnt);
  const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
  uint32_t base = 0;
  if (lane == 0 && total) base = atomicAdd(&lb.fill[part], total);
  base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
  const bool fits = base + total <= lb.part_cap;
  // the wave's lines, in order, for the binning wave of the same index
  if (lane == 0) lb.wave_lines[wave] = make_uint2(part * lb.part_cap + base, fits ? total : 0u);
  if (total == 0) return;
  if (!fits) {
    if (lane == 0) {
      atomicOr(&lb.st->overflow, 4u);
      atomicMax(&lb.st->n_line_part, base + total);
    }
    return;
  }
  if (!cnt) return;
  size_t o = (size_t)part * lb.part_cap + base + incl - cnt;
  if (cnt <= (uint32_t)kLineKeep && chain) {
    for (uint32_t k = 0; k < cnt; ++k, ++o) {
      const double2 a = keep[k], c = keep[k + 1];
      lb.lines[o] = make_double4(a.x, a.y, c.x, c.y);
      if (MULTI) lb.line_path[o] = pid;
    }
    return;
  }
  // more lines than were kept: walk the subtree again
  auto put = [&](Pt a, Pt c) {
    if (MULTI) lb.line_path[o] = pid;
    lb.lines[o++] = make_double4(a.x, a.y, c.x, c.y);
  };
  if (leaf) {
    lines_of(n.p[0], n.p[NP - 1], put);
    return;
  }
  rzg::traverse(n, fl, [&](Pt a, Pt c) { lines_of(a, c, put); });
}

// The flatten half of k_flat_bin: flat_item's count and placement, but the
// lines of a thread with at most kLineKeep of them stay in `keep` (LDS) for
// the count pass of the same wave; only longer runs go to the line buffer.
struct FlatRes {
  uint32_t cnt;    // this lane's lines
  uint32_t pre;    // lines of the lanes before it
  uint32_t total;  // the wave's lines (0 when none, or they did not fit)
  uint32_t line0;  // the wave's first line index
  uint32_t pid;    // batched: the lane's path
  bool kept;       // the lane's lines are in its LDS keep
};
template <int NP, bool MULTI, bool AE>
__device__ void flat_keep(const DPath& p, uint32_t slot, uint32_t j, int D0, bool active,
                          const Xform& x, double fl, const PathInfo* pi, const uint32_t* slot_path,
                          const LineBufs& lb, double2* keep, uint32_t part, uint32_t wave,
                          FlatRes* fr) {
  rzg::Node<NP> n;
  bool walk, leaf;
  int err;
  uint32_t pid;
  flat_subtree<NP, MULTI>(p, slot, j, D0, active, x, fl, pi, slot_path, &n, &walk, &leaf, &err, &pid);
  double ae_h = 0.0;
  if (AE) ae_h = MULTI ? (double)(active ? pi[pid].g.H : 0) : lb.ae_h;
  const rzg::WrapClip wq = wrap_clip<MULTI>(lb, pi, pid, active);
  bool chain = true;  // as in flat_item
  int werr = rzg::ST_OK;
  auto lines_of = [&](Pt a, Pt c, auto&& f) {
    if (AE) {
      rzg::ae_chain(a, c, ae_h, f);
    } else if (rzg::wraps(a, c, wq)) {
      chain = false;
      if (rzg::wrap_lines(a, c, wq, [&](Pt u, Pt v) { f(u, rzg::mkpt(rzg::wrap_tag(), v.y)); }))
        werr = rzg::ST_WRAP;
    } else {
      f(a, c);
    }
  };
  uint32_t cnt = 0;
  auto count = [&](Pt a, Pt c) {
    if (cnt < (uint32_t)kLineKeep) {
      if (cnt == 0) keep[0] = make_double2(a.x, a.y);
      keep[cnt + 1] = make_double2(c.x, c.y);
    }
    ++cnt;
  };
  if (leaf) lines_of(n.p[0], n.p[NP - 1], count);
  if (walk) {
    const int r = rzg::traverse(n, fl, [&](Pt a, Pt c) { lines_of(a, c, count); });
    if (r != rzg::ST_OK) err = r;
  }
  if (err == rzg::ST_OK) err = werr;
  if (err != rzg::ST_OK) {
    set_err(lb.st, (uint32_t)err);
    cnt = 0;
  }
  const int lane = threadIdx.x & 63;
  const uint32_t incl = (uint32_t)dpp_add_scan((int)cnt);
  const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
  uint32_t base = 0;
  if (lane == 0 && total) base = atomicAdd(&lb.fill[part], total);
  base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
  const bool fits = base + total <= lb.part_cap;
  fr->cnt = cnt;
  fr->kept = cnt <= (uint32_t)kLineKeep && chain;
  fr->pre = incl - cnt;
  fr->total = fits ? total : 0u;
  fr->line0 = part * lb.part_cap + base;
  fr->pid = pid;
  // the wave's line range, as k_flat_lines records it (the scatter reads it)
  if (lane == 0) lb.wave_lines[wave] = make_uint2(fr->line0, fr->total);
  if (total == 0) return;
  if (!fits) {
    if (lane == 0) {
      atomicOr(&lb.st->overflow, 4u);
      atomicMax(&lb.st->n_line_part, base + total);
    }
    return;
  }
  if (fr->kept) return;
  // more lines than were kept: walk the subtree again into the line buffer
  size_t o = (size_t)fr->line0 + incl - cnt;
  auto put = [&](Pt a, Pt c) { lb.lines[o++] = make_double4(a.x, a.y, c.x, c.y); };
  if (leaf) {
    lines_of(n.p[0], n.p[NP - 1], put);
    return;
  }
  rzg::traverse(n, fl, [&](Pt a, Pt c) { lines_of(a, c, put); });
}

// Blocks [0, bl) handle line slots, [bl, bl + bq) quad virtual items, the rest
// cubic virtual items. Every lane of a wave reaches the wave operations.
#define RZ_FLAT_DISPATCH(ITEM, ...)                                                        \
  const uint32_t blk = blockIdx.x / (kBlock / kFlatBlock),                                 \
                 t = (blockIdx.x % (kBlock / kFlatBlock)) * kFlatBlock + threadIdx.x;      \
  if (blk < bl) {                                                                          \
    const uint32_t slot = p.order[blk * kBlock + t];                                       \
    ITEM<2, MULTI, AE, WRAP>(p, slot, 0, 0, slot != ~0u, x, fl, pi, slot_path, __VA_ARGS__);     \
  } else if (blk < bl + bq) {                                                              \
    const uint32_t v = (blk - bl) * kBlock + t;                                            \
    const uint32_t slot = p.order[p.blocks_line * kBlock + (v >> dq)];                     \
    ITEM<3, MULTI, AE, WRAP>(p, slot, v & ((1u << dq) - 1u), dq, slot != ~0u, x, fl, pi,        \
                       slot_path, __VA_ARGS__);                                            \
  } else {                                                                                 \
    const uint32_t v = (blk - bl - bq) * kBlock + t;                                       \
    const uint32_t slot = p.order[(p.blocks_line + p.blocks_quad) * kBlock + (v >> dc)];   \
    ITEM<4, MULTI, AE, WRAP>(p, slot, v & ((1u << dc) - 1u), dc, slot != ~0u, x, fl, pi,        \
                       slot_path, __VA_ARGS__);                                            \
  }

// Workgroup size of the flatten: the slot blocks above are kBlock threads,
// each run as kBlock / kFlatBlock workgroups. One wave per workgroup: a wave
// with long subtrees then holds only its own LDS and wave slot.
constexpr uint32_t kFlatBlock = 64;

template <bool MULTI, bool AE, bool WRAP>
__global__ __launch_bounds__(kFlatBlock) void k_flat_lines(DPath p, Xform x, double fl, uint32_t bl,
                                                            uint32_t bq, int dq, int dc,
                                                            const PathInfo* pi,
                                                            const uint32_t* slot_path, LineBufs lb) {
  __shared__ double2 keep_all[kFlatBlock * (kLineKeep + 1)];
  if (lb.st->err || lb.st->overflow) return;  // e.g. a batched scene whose slices did not fit
  double2* keep = keep_all + threadIdx.x * (kLineKeep + 1);
  const uint32_t wave = blockIdx.x * (kFlatBlock / 64) + (threadIdx.x >> 6);
  const uint32_t part = wave % kLineParts;
  RZ_FLAT_DISPATCH(flat_item, lb, keep, part, wave)
}


// ---------------------------------------------------------------------------
// Binning of lines into cell records.
//   k_bin_lines   the count pass: band-bottom crossings, right-edge rows,
//                 events, and every cell record with its place among its
//                 tile's records, which is only known up to one offset per
//                 (wave, tile): a wave counts its records per tile in LDS and
//                 adds each count to the tile's total with one atomic at its
//                 end, which returns that offset;
//   (scans of the tile and band counts)
//   k_bin_scatter writes each listed record to tile offset + wave offset +
//                 rank (no atomics), and moves the listed events into their
//                 bands' slots.
// Each wave lists its records in a region of its own; a wave that fills it
// continues in an overflow list of kParts parts, each with its own bump
// counter, taking kChunk entries at a time and padding what it leaves unused
// with tile ~0u (the scatter skips padding).
// ---------------------------------------------------------------------------
constexpr uint32_t kParts = 16;
constexpr uint32_t kChunk = 128;
// Per-wave LDS map tile -> record count (open addressing). A tile holds
// fewer than 2^24 records.
constexpr int kMapBits = 7;
constexpr uint32_t kMap = 1u << kMapBits;
constexpr int kMapProbe = 16;
constexpr uint32_t kMapNone = 0xffu;  // record ranked in its tile directly
constexpr uint32_t kEmpty = ~0u;

struct BinBufs {
  uint32_t* tile_count;   // [tiles] records per tile
  const uint32_t* tile_off;  // [tiles] their exclusive prefix (after the scan)
  int32_t* bd_diff;       // [tiles] band backdrop differences
  uint32_t* band_ev;      // counts per band; after the scan: cursors
  float4* recs;
  uint32_t rec_cap;
  int4* events;           // (band, col, yrel bits, sign bits)
  uint32_t ev_cap;
  long long* redge;       // [H + 2] right-edge row differences (int64: a row may be crossed
                          // by any number of lines beyond the clip)
  // record list: two entries per line slot (each binning wave's region is
  // those of its lines), then [kParts][part_cap] overflow parts
  uint2* list_meta;       // tile (~0u: padding), map slot << 24 | rank within the wave's
                          // records of the tile
  float4* list_rec;       // the record
  uint32_t* ovf_wave;     // [kParts][part_cap] wave of each overflow entry
  uint32_t n_waves, part_cap;
  size_t list_regions;    // entries before the overflow parts
  const uint2* wave_lines;  // per flatten wave: first line, line count (one binning wave each)
  uint32_t* wave_used;    // [n_waves] entries of each wave's region
  uint32_t* wave_map;     // [n_waves][kMap] the wave's offset in each mapped tile
  uint32_t* fill;         // [kParts + 1] entries taken per overflow part, then listed events
  int4* ev_list;          // [ev_cap] events of the count pass
  uint32_t* ev_band;      // [ev_cap] their band slots (band base + band)
  Status* st;
};

// Count (PASS 0) or place (PASS 1) the free-endpoint event of vertex p.
template <int PASS>
__device__ inline void put_event(const BinBufs& b, const rzg::BinGrid& g, uint32_t band_base, Pt p,
                                 float sign) {
  rzg::Event ev;
  if (!rzg::make_event(p, sign, g, &ev)) return;
  if (PASS == 0) {
    atomicAdd(&b.band_ev[band_base + ev.band], 1u);
  } else {
    uint32_t slot = atomicAdd(&b.band_ev[band_base + ev.band], 1u);
    if (slot < b.ev_cap)
      b.events[slot] = make_int4(ev.band, ev.col, __float_as_int(ev.yrel), __float_as_int(ev.sign));
    else atomicOr(&b.st->overflow, 2u);
  }
}

__device__ inline int4 event_bits(const rzg::Event& ev) {
  return make_int4(ev.band, ev.col, __float_as_int(ev.yrel), __float_as_int(ev.sign));
}

// Count (per band) and list up to two events per lane: one atomic per wave
// takes the list entries. Every lane of the wave takes part.
__device__ inline void list_events(const BinBufs& b, int nev, const int4& ev0, uint32_t evb0,
                                   const int4& ev1, uint32_t evb1, int lane) {
  const uint64_t below = (1ull << lane) - 1ull;
  if (nev >= 1) atomicAdd(&b.band_ev[evb0], 1u);
  if (nev == 2) atomicAdd(&b.band_ev[evb1], 1u);
  const uint64_t e1 = __ballot(nev >= 1), e2 = __ballot(nev == 2);
  if (!e1) return;
  const uint32_t etot = (uint32_t)(__popcll(e1) + __popcll(e2));
  uint32_t ebase = 0;
  if (lane == 0) ebase = atomicAdd(&b.fill[kParts], etot);
  ebase = (uint32_t)__shfl((int)ebase, 0, 64);
  const uint32_t k = ebase + (uint32_t)(__popcll(e1 & below) + __popcll(e2 & below));
  if (ebase + etot > b.ev_cap) {
    if (lane == 0) atomicOr(&b.st->overflow, 2u);
    return;
  }
  if (nev >= 1) {
    b.ev_list[k] = ev0;
    b.ev_band[k] = evb0;
  }
  if (nev == 2) {
    b.ev_list[k + 1] = ev1;
    b.ev_band[k + 1] = evb1;
  }
}

// Exclusive prefix sum of v over the wave (every lane takes part); the wave
// total goes to *total.
__device__ inline uint32_t wave_excl_scan(uint32_t v, int lane, uint32_t* total) {
  const uint32_t inc = (uint32_t)dpp_add_scan((int)v);
  *total = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
  return inc - v;
}

// Order this wave's LDS writes before its following LDS reads (LDS executes a
// wave's operations in issue order; this keeps the compiler from moving them).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Owner lane of task j0 + lane, lane o owning tasks [pre, pre + n). Each lane
// whose range starts in [j0, j0 + 64) marks its start in the wave's 64-entry
// LDS `marks`; the owner is the running maximum of the marks, or `carry`, the
// owner of task j0 - 1 (-1 before the first round), which is updated. Every
// lane takes part.
__device__ inline int task_owner(int* marks, uint32_t pre, uint32_t n, uint32_t j0, int lane,
                                 int* carry) {
  marks[lane] = -1;
  wave_lds_sync();
  if (n > 0 && pre >= j0 && pre - j0 < 64u) marks[pre - j0] = lane;
  wave_lds_sync();
  const int o = max(dpp_max_scan(marks[lane]), *carry);
  *carry = __builtin_amdgcn_readlane(o, 63);
  return o;
}

// Slot of `tile` in a wave's LDS map, inserting it if absent; -1 when the
// probe sequence is full.
__device__ inline int map_slot(uint32_t* keys, uint32_t tile) {
  uint32_t h = (tile * 2654435761u) >> (32 - kMapBits);
  for (int step = 0; step < kMapProbe; ++step) {
    const uint32_t k = __atomic_load_n(&keys[h], __ATOMIC_RELAXED);
    if (k == tile) return (int)h;
    if (k == kEmpty) {
      const uint32_t prev = atomicCAS(&keys[h], kEmpty, tile);
      if (prev == kEmpty || prev == tile) return (int)h;
    }
    h = (h + 1) & (kMap - 1);
  }
  return -1;
}

// A line's effective polyline and its path's grid, staged in LDS for the band
// tasks of its wave.
struct LineStage {
  double px[3], py[3];
  int ty0[2];  // first band of each piece
  int nb0;     // bands of piece 0
  int H, tiles_x, phx, phy;
  uint32_t tb;  // the path's first tile
};

// A band part, staged for the tile tasks.
struct BandStage {
  double sx, sy, ex, ey, r0;
  int ty, tx0;
  int line;  // lane of its line
  int pad;
};

// The part of a line's grid the band and tile steps use.
__device__ inline rzg::BinGrid stage_grid(const LineStage& S) {
  rzg::BinGrid g{};
  g.H = S.H;
  g.tiles_x = S.tiles_x;
  g.phx = S.phx;
  g.phy = S.phy;
  return g;
}

// A gap between the previous segment's end `a` and a segment's start `b`
// (arc endpoints are recomputed by sin/cos and miss their neighbours by
// ~1e-13): the chain has a free end at a (+V) and a free start at b (-V).
// Both events are made here; a pair that lands on the same band, column and
// f32 row offset contributes exactly zero and is dropped. Ends created by the
// right clip are reported by the clip itself. The events are counted and
// listed like the lines' (k_bin_scatter places them). One lane per gap slot
// (i = ~0u: none); every lane of the wave takes part.
struct GapArgs {
  DPath p;
  Xform x;
  const uint32_t* slot_path;  // batched scenes and atlases: path of each slot
};

template <bool MULTI>
__device__ void gap_events(const GapArgs& ga, const rzg::BinGrid& g0, const BinBufs& b,
                           const PathInfo* pi, uint32_t i, int lane) {
  int nev = 0;
  int4 ev0{}, ev1{};
  uint32_t evb0 = 0, evb1 = 0;
  if (i != ~0u) {
    const uint32_t slot = ga.p.gap_slots[i];
    rzg::BinGrid g = g0;
    const double* m = ga.x.m;
    uint32_t band_base = 0;
    bool live = true;
    if (MULTI) {
      const PathInfo& P = pi[ga.slot_path[slot]];
      live = P.live != 0;
      g = P.g;
      m = P.m;
      band_base = P.band_base;
    }
    if (live) {
      const uint32_t k = ga.p.slot_item[slot];  // a segment (closing lines never start a gap)
      const double* qa = ga.p.pts + 2 * (size_t)(ga.p.pt_off[k] - 1);  // previous segment's last point
      const double* qb = ga.p.pts + 2 * (size_t)ga.p.pt_off[k];
      const Pt a = rzg::tr_apply(m, rzg::mkpt(qa[0], qa[1]));
      const Pt bb = rzg::tr_apply(m, rzg::mkpt(qb[0], qb[1]));
      const bool clip = g.mode != rzg::EDGE_NONE;
      const bool ea = !(clip && a.x > g.wc), eb = !(clip && bb.x > g.wc);
      rzg::Event va, vb;
      const Pt ap = rzg::mkpt(fmax(a.x, 0.0), a.y), bp = rzg::mkpt(fmax(bb.x, 0.0), bb.y);
      const bool ha = ea && rzg::make_event(ap, 1.0f, g, &va);
      const bool hb = eb && rzg::make_event(bp, -1.0f, g, &vb);
      if (!(ha && hb && va.band == vb.band && va.col == vb.col && va.yrel == vb.yrel)) {
        if (ha) {
          ev0 = event_bits(va);
          evb0 = band_base + (uint32_t)va.band;
          nev = 1;
        }
        if (hb) {
          (nev ? ev1 : ev0) = event_bits(vb);
          (nev ? evb1 : evb0) = band_base + (uint32_t)vb.band;
          ++nev;
        }
      }
    }
  }
  list_events(b, nev, ev0, evb0, ev1, evb1, lane);
}

// Count pass. A wave takes the lines of the flatten wave of the same index
// (consecutive along the path), 64 at a time, one per lane: right clip and
// left split, right-edge column, events, and the band range of each piece.
// The (line, band) pairs of the 64 lines are then dealt out to the lanes 64
// at a time, and the (band part, tile column) pairs of each such round
// likewise, so a line crossing many cells keeps many lanes busy rather than
// one. Each record takes its rank from its tile's count in the wave's LDS
// map (one LDS atomic) and goes to the wave's region of the record list.
// Lines beyond the right clip only add to the right-edge column; two such
// lines on neighbouring lanes that continue each other drop their shared
// endpoint (its two terms cancel exactly). Waves past the flatten waves take
// the gap slots.

// Workgroup size of the count pass. Its waves are independent and their
// work varies with their lines, so one wave per workgroup.
constexpr uint32_t kBinBlock = 64;

// The LDS of one count-pass wave.
struct BinLDS {
  LineStage lstage[64];
  BandStage bstage[64];
  uint32_t keys[kMap];
  uint32_t cnts[kMap];
  int marks[64];
};

// The count pass of one wave over its lines: wl_range = (first line index,
// line count), `load(c, &line, &path)` sets this lane's line c + lane (and,
// batched, its path) where c + lane < the count and leaves them otherwise;
// it is called for c = 0, 64, 128, ... in that order, by every lane.
template <bool MULTI, class Load>
__device__ inline void bin_wave(BinLDS& lds, uint32_t wave, uint2 wl_range, const rzg::BinGrid& g0,
                                const PathInfo* pi, const BinBufs& b, Load&& load) {
  LineStage* wl = lds.lstage;
  BandStage* wb = lds.bstage;
  uint32_t* keys = lds.keys;
  uint32_t* cnts = lds.cnts;
  int* marks = lds.marks;
  const int lane = threadIdx.x & 63;
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t q = (uint32_t)lane; q < kMap; q += 64) {
    keys[q] = kEmpty;
    cnts[q] = 0;
  }
  const uint32_t part = wave % kParts;
  const size_t region = 2 * (size_t)wl_range.x;
  const uint32_t region_cap = 2 * wl_range.y;
  const size_t obase = b.list_regions + (size_t)part * b.part_cap;
  uint32_t used = 0;                   // entries of the wave's region
  uint32_t cbase = 0, cused = kChunk;  // the wave's current overflow chunk (none yet)
  bool lover = false;                  // the overflow part is full (the run is re-done)

  // Rank and list one record per lane (has = false: none); every lane takes
  // part. Each record takes its rank from its tile's count in the wave's LDS
  // map (lanes with equal tiles are serialized by the LDS atomic).
  auto emit = [&](bool has, int tile, const rzg::Rec& r) {
    uint32_t rank = 0;
    if (has) {
      const int h = map_slot(keys, (uint32_t)tile);
      if (h >= 0) {
        rank = ((uint32_t)h << 24) | atomicAdd(&cnts[h], 1u);
      } else {
        // the map is full: rank directly among all of the tile's records
        rank = (kMapNone << 24) | atomicAdd(&b.tile_count[tile], 1u);
      }
    }
    // append to the wave's region, or past its end to the overflow list
    const uint64_t hm = __ballot(has);
    const uint32_t cnt = (uint32_t)__popcll(hm);
    const uint32_t lrank = (uint32_t)__popcll(hm & below);
    if (cnt && used + cnt <= region_cap) {
      if (has) {
        const size_t q = region + used + lrank;
        b.list_meta[q] = make_uint2((uint32_t)tile, rank);
        b.list_rec[q] = make_float4(r.x0, r.y0, r.x1, r.y1);
      }
      used += cnt;
    } else if (cnt) {
      if (cused + cnt > kChunk) {
        // pad the rest of the chunk and take a new one
        if (!lover)
          for (uint32_t q = cused + lane; q < kChunk; q += 64)
            b.list_meta[obase + cbase + q] = make_uint2(~0u, 0u);
        uint32_t nb = 0;
        if (lane == 0) nb = atomicAdd(&b.fill[part], kChunk);
        nb = (uint32_t)__shfl((int)nb, 0, 64);
        if (nb + kChunk > b.part_cap) {
          if (lane == 0) {
            atomicOr(&b.st->overflow, 16u);
            atomicMax(&b.st->n_list, nb + kChunk);
          }
          lover = true;
        }
        cbase = nb;
        cused = 0;
      }
      if (has && !lover) {
        const size_t q = obase + cbase + cused + lrank;
        b.list_meta[q] = make_uint2((uint32_t)tile, rank);
        b.list_rec[q] = make_float4(r.x0, r.y0, r.x1, r.y1);
        b.ovf_wave[q - b.list_regions] = wave;
      }
      cused += cnt;
    }
  };

  // each group's line (and path) is loaded while the previous group is binned
  double4 fL = make_double4(0, 0, 0, 0);
  uint32_t fpid = 0;
  auto fetch = [&](uint32_t c) { load(c, &fL, &fpid); };
  // batched forms: the group's path data, loaded during the previous group
  rzg::BinGrid fg = g0;
  uint32_t ftb = 0, fbb = 0, frb0 = 0;
  auto fetch_path = [&](bool h) {
    if (MULTI && h) {
      const PathInfo& P = pi[fpid];
      fg = P.g;
      ftb = P.tile_base;
      fbb = P.band_base;
      frb0 = P.redge_base;
    }
  };
  fetch(0);
  fetch_path((uint32_t)lane < wl_range.y);
  for (uint32_t c0 = 0; c0 < wl_range.y; c0 += 64) {
    bool have = c0 + (uint32_t)lane < wl_range.y;
    const double4 L = have ? fL : make_double4(0, 0, 0, 0);
    uint32_t pid = have ? fpid : 0u;
    const bool next = c0 + 64 < wl_range.y;
    if (next) fetch(c0 + 64);
    uint32_t tb = 0, bb = 0, rb0 = 0;
    rzg::BinGrid g = g0;
    if (MULTI && have) {
      g = fg;
      tb = ftb;
      bb = fbb;
      rb0 = frb0;
    }
    // a line of rzg::wrap_lines: vertical on L.x, both ends free
    const bool wrapv = have && rzg::is_wrap_tag(L.z);
    const Pt p0 = rzg::mkpt(L.x, L.y), p1 = rzg::mkpt(wrapv ? L.x : L.z, L.w);
    if (have && !(isfinite(p0.x) && isfinite(p0.y) && isfinite(p1.x) && isfinite(p1.y))) {
      set_err(b.st, rzg::ST_NAN);
      have = false;
    }
    const bool proj = have && g.mode != rzg::EDGE_NONE && p0.x > g.wc && p1.x > g.wc;
    {
      const double nx = lane_next(L.x), ny = lane_next(L.y);
      const double px = lane_prev(L.z), py = lane_prev(L.w);
      const bool nproj = lane_next((int)proj) != 0;
      const bool pproj = lane_prev((int)proj) != 0;
      const uint32_t npid = (uint32_t)lane_next((int)pid), ppid = (uint32_t)lane_prev((int)pid);
      if (proj) {
        const bool skip_end = lane < 63 && nproj && npid == pid && nx == p1.x && ny == p1.y;
        const bool skip_start = lane > 0 && pproj && ppid == pid && px == p0.x && py == p0.y;
        auto add = [&](int k, int v) {
          atomicAdd((unsigned long long*)&b.redge[rb0 + k], (unsigned long long)(long long)v);
        };
        if (!skip_end) rzg::redge_endpoint(p1.y, 1, g, add);
        if (!skip_start) rzg::redge_endpoint(p0.y, -1, g, add);
      }
    }
    int nev = 0;
    int4 ev0{}, ev1{};
    uint32_t evb0 = 0, evb1 = 0;
    int nb0 = 0, nb1 = 0;
    if (have && !proj) {
      rzg::EffLine e = rzg::effective_line(p0, p1, g);
      if (wrapv) e.clip0 = e.clip1 = true;
      rzg::Event ev;
      if (e.clip0 && rzg::make_event(e.p[0], -1.0f, g, &ev)) {
        ev0 = event_bits(ev);
        evb0 = bb + (uint32_t)ev.band;
        nev = 1;
      }
      if (e.clip1 && rzg::make_event(e.n == 2 ? e.p[2] : e.p[1], 1.0f, g, &ev)) {
        (nev ? ev1 : ev0) = event_bits(ev);
        (nev ? evb1 : evb0) = bb + (uint32_t)ev.band;
        ++nev;
      }
      int ty0 = 0, ty1 = -1, uy0 = 0, uy1 = -1;
      if (rzg::piece_bands(e.p[0], e.p[1], g, &ty0, &ty1) && ty1 >= ty0) nb0 = ty1 - ty0 + 1;
      if (e.n == 2 && rzg::piece_bands(e.p[1], e.p[2], g, &uy0, &uy1) && uy1 >= uy0)
        nb1 = uy1 - uy0 + 1;
      LineStage& S = wl[lane];
      S.px[0] = e.p[0].x;
      S.py[0] = e.p[0].y;
      S.px[1] = e.p[1].x;
      S.py[1] = e.p[1].y;
      S.px[2] = e.p[2].x;
      S.py[2] = e.p[2].y;
      S.ty0[0] = ty0;
      S.ty0[1] = uy0;
      S.nb0 = nb0;
      S.H = g.H;
      S.tiles_x = g.tiles_x;
      S.phx = g.phx;
      S.phy = g.phy;
      S.tb = tb;
    }
    list_events(b, nev, ev0, evb0, ev1, evb1, lane);
    if (MULTI && next) fetch_path(c0 + 64 + (uint32_t)lane < wl_range.y);
    // band tasks, 64 per round
    uint32_t n_band;
    const uint32_t pre1 = wave_excl_scan((uint32_t)(nb0 + nb1), lane, &n_band);
    wave_lds_sync();
    int carry1 = -1;
    for (uint32_t j0 = 0; j0 < n_band; j0 += 64) {
      const uint32_t j = j0 + lane;
      const int o = task_owner(marks, pre1, (uint32_t)(nb0 + nb1), j0, lane, &carry1);
      const uint32_t opre = (uint32_t)__shfl((int)pre1, o, 64);
      int nt = 0;
      if (j < n_band) {
        const LineStage& S = wl[o];
        const int k = (int)(j - opre);
        const int pc = k < S.nb0 ? 0 : 1;
        const int ty = pc == 0 ? S.ty0[0] + k : S.ty0[1] + (k - S.nb0);
        const Pt a = rzg::mkpt(S.px[pc], S.py[pc]), c = rzg::mkpt(S.px[pc + 1], S.py[pc + 1]);
        const rzg::BinGrid gt = stage_grid(S);
        rzg::BandPiece bp;
        rzg::band_piece(a, c, gt, ty, &bp);
        if (bp.cross_col >= 0)
          atomicAdd(&b.bd_diff[S.tb + ty * gt.tiles_x + bp.cross_col], c.y > a.y ? 1 : -1);
        if (bp.any) {
          nt = bp.tx1 - bp.tx0 + 1;
          BandStage& B = wb[lane];
          B.sx = bp.s.x;
          B.sy = bp.s.y;
          B.ex = bp.e.x;
          B.ey = bp.e.y;
          B.r0 = bp.r0;
          B.ty = ty;
          B.tx0 = bp.tx0;
          B.line = o;
        }
      }
      // tile tasks of this round's band parts, 64 per round
      uint32_t n_tile;
      const uint32_t pre2 = wave_excl_scan((uint32_t)nt, lane, &n_tile);
      wave_lds_sync();
      int carry2 = -1;
      for (uint32_t k0 = 0; k0 < n_tile; k0 += 64) {
        const uint32_t k = k0 + lane;
        const int o2 = task_owner(marks, pre2, (uint32_t)nt, k0, lane, &carry2);
        const uint32_t opre2 = (uint32_t)__shfl((int)pre2, o2, 64);
        bool has = false;
        int tile = -1;
        rzg::Rec r{};
        if (k < n_tile) {
          const BandStage& B = wb[o2];
          const LineStage& S = wl[B.line];
          const rzg::BinGrid gt = stage_grid(S);
          const int tx = B.tx0 + (int)(k - opre2);
          has = rzg::tile_record(rzg::mkpt(B.sx, B.sy), rzg::mkpt(B.ex, B.ey), B.r0, tx, gt, &r);
          tile = (int)S.tb + B.ty * S.tiles_x + tx;
        }
        emit(has, tile, r);
      }
      wave_lds_sync();  // the next band round re-stages the band parts
    }
    wave_lds_sync();  // the next lines re-stage the polylines
  }
  if (!lover && cused < kChunk)
    for (uint32_t q = cused + lane; q < kChunk; q += 64) b.list_meta[obase + cbase + q] = make_uint2(~0u, 0u);
  // add the wave's counts to the tiles' totals; each atomic returns the
  // wave's offset among the tile's records
  wave_lds_sync();
  for (uint32_t q = (uint32_t)lane; q < kMap; q += 64) {
    const uint32_t t = keys[q];
    if (t != kEmpty) b.wave_map[(size_t)wave * kMap + q] = atomicAdd(&b.tile_count[t], cnts[q]);
  }
  if (lane == 0) b.wave_used[wave] = used;
}

template <bool MULTI>
__global__ __launch_bounds__(kBinBlock) void k_bin_lines(const double4* lines,
                                                          const uint32_t* line_path, rzg::BinGrid g0,
                                                          const PathInfo* pi, BinBufs b, GapArgs ga) {
  __shared__ BinLDS S;
  if (b.st->err || b.st->overflow) return;
  const int lane = threadIdx.x & 63;
  const uint32_t wave = blockIdx.x * (kBinBlock / 64) + (threadIdx.x >> 6);
  if (wave >= b.n_waves) {
    // the waves past the lines' take the gap slots, 64 each
    const uint32_t i = (wave - b.n_waves) * 64 + (uint32_t)lane;
    gap_events<MULTI>(ga, g0, b, pi, i < ga.p.n_gaps ? i : ~0u, lane);
    return;
  }
  const uint2 wl_range = b.wave_lines[wave];
  if (wl_range.y == 0) return;  // the scatter skips the same regions
  bin_wave<MULTI>(S, wave, wl_range, g0, pi, b, [&](uint32_t c, double4* L, uint32_t* pid) {
    const size_t i = (size_t)wl_range.x + c + (uint32_t)lane;
    if (c + (uint32_t)lane < wl_range.y) {
      *L = lines[i];
      if (MULTI) *pid = line_path[i];
    }
  });
}

// Flatten and count pass in one kernel: each wave flattens its items (as
// k_flat_lines) and bins the lines it made (as k_bin_lines), reading them
// from its own LDS rather than back from HBM; a thread's run of more than
// kLineKeep lines goes through the line buffer as before. Waves past the
// flatten waves take the gap slots. Bit for bit the same records, events,
// backdrops and right-edge rows as the two kernels.
template <bool MULTI, bool AE>
__global__ __launch_bounds__(64) void k_flat_bin(DPath p, Xform x, double fl, uint32_t bl,
                                                 uint32_t bq, int dq, int dc, const PathInfo* pi,
                                                 const uint32_t* slot_path, LineBufs lb,
                                                 rzg::BinGrid g0, BinBufs b, GapArgs ga) {
  __shared__ double2 keep_all[64 * (kLineKeep + 1)];
  __shared__ BinLDS S;
  if (b.st->err || b.st->overflow) return;
  const int lane = threadIdx.x;
  const uint32_t wave = blockIdx.x;
  if (wave >= lb.n_waves) {
    const uint32_t i = (wave - lb.n_waves) * 64 + (uint32_t)lane;
    gap_events<MULTI>(ga, g0, b, pi, i < ga.p.n_gaps ? i : ~0u, lane);
    return;
  }
  const uint32_t part = wave % kLineParts;
  FlatRes fr;
  {
    double2* keep = keep_all + lane * (kLineKeep + 1);
    const uint32_t blk = wave / (kBlock / 64), t = (wave % (kBlock / 64)) * 64 + (uint32_t)lane;
    if (blk < bl) {
      const uint32_t slot = p.order[blk * kBlock + t];
      flat_keep<2, MULTI, AE>(p, slot, 0, 0, slot != ~0u, x, fl, pi, slot_path, lb, keep, part,
                              wave, &fr);
    } else if (blk < bl + bq) {
      const uint32_t v = (blk - bl) * kBlock + t;
      const uint32_t slot = p.order[p.blocks_line * kBlock + (v >> dq)];
      flat_keep<3, MULTI, AE>(p, slot, v & ((1u << dq) - 1u), dq, slot != ~0u, x, fl, pi,
                              slot_path, lb, keep, part, wave, &fr);
    } else {
      const uint32_t v = (blk - bl - bq) * kBlock + t;
      const uint32_t slot = p.order[(p.blocks_line + p.blocks_quad) * kBlock + (v >> dc)];
      flat_keep<4, MULTI, AE>(p, slot, v & ((1u << dc) - 1u), dc, slot != ~0u, x, fl, pi,
                              slot_path, lb, keep, part, wave, &fr);
    }
  }
  if (fr.total == 0) return;
  // the kept lines (LDS) and the long runs (line buffer) written by other
  // lanes are read below
  __syncthreads();
  int carry = -1;  // owner lane of the line before the current group
  bin_wave<MULTI>(S, wave, make_uint2(fr.line0, fr.total), g0, pi, b,
                  [&](uint32_t c, double4* L, uint32_t* pid) {
    const int o = task_owner(S.marks, fr.pre, fr.cnt, c, lane, &carry);
    const uint32_t opre = (uint32_t)__shfl((int)fr.pre, o, 64);
    const uint32_t opid = (uint32_t)__shfl((int)fr.pid, o, 64);
    const bool okept = __shfl((int)fr.kept, o, 64) != 0;
    const uint32_t i = c + (uint32_t)lane;
    if (i < fr.total) {
      if (okept) {
        const double2* ok = keep_all + o * (kLineKeep + 1) + (i - opre);
        const double2 a = ok[0], e = ok[1];
        *L = make_double4(a.x, a.y, e.x, e.y);
      } else {
        *L = lb.lines[(size_t)fr.line0 + i];
      }
      if (MULTI) *pid = opid;
    }
  });
}

// Write one listed record to its slot: tile offset + the wave's offset in the
// tile (through its map slot, kMapNone: already included) + rank.
__device__ inline void place_record(const BinBufs& b, uint32_t tile, uint32_t rank, uint32_t wave,
                                    const float4& r) {
  const uint32_t h = rank >> 24;
  uint32_t slot = b.tile_off[tile] + (rank & 0xffffffu);
  if (h != kMapNone) slot += b.wave_map[(size_t)wave * kMap + h];
  if (slot < b.rec_cap) b.recs[slot] = r;
  else atomicOr(&b.st->overflow, 1u);
}

// Place the listed records: one wave per count-pass region (grid-stride over
// the regions of the waves that had lines), then the overflow parts one entry
// per thread.
constexpr uint32_t kScatterBlock = 256;
constexpr uint32_t kScatterWaves = 32768;  // waves of the scatter grid at most

// Place the events the count pass listed into their bands' slots (blocks
// from `first` on; the first of them also totals the lines, for reporting).
__device__ void place_events(const BinBufs& b, const uint32_t* line_fill, uint32_t first) {
  const uint32_t blk = blockIdx.x - first, nblk = gridDim.x - first;
  if (blk == 0 && threadIdx.x < 64) {
    uint32_t v = line_fill[threadIdx.x];
    for (int o = 32; o >= 1; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, 64);
    if (threadIdx.x == 0) b.st->n_lines = v;
  }
  if (b.st->err || b.st->overflow) return;
  const uint32_t n = b.fill[kParts];
  for (uint32_t k = blk * kScatterBlock + threadIdx.x; k < n; k += nblk * kScatterBlock) {
    const uint32_t slot = atomicAdd(&b.band_ev[b.ev_band[k]], 1u);
    if (slot < b.ev_cap) b.events[slot] = b.ev_list[k];
    else atomicOr(&b.st->overflow, 2u);
  }
}

// Blocks [0, n_rec_blocks) place the listed records, the rest the events.
__global__ __launch_bounds__(kScatterBlock) void k_bin_scatter(BinBufs b, const uint32_t* line_fill,
                                                               uint32_t n_rec_blocks) {
  if (blockIdx.x >= n_rec_blocks) {
    place_events(b, line_fill, n_rec_blocks);
    return;
  }
  if (b.st->err || b.st->overflow) return;
  const uint32_t wpb = kScatterBlock / 64, wstride = n_rec_blocks * wpb;
  // the next region's line range and entry count are loaded while the
  // current region's entries are placed
  uint32_t w = blockIdx.x * wpb + (threadIdx.x >> 6);
  uint2 r = make_uint2(0u, 0u);
  uint32_t used = 0;
  if (w < b.n_waves) {
    r = b.wave_lines[w];
    used = b.wave_used[w];
  }
  while (w < b.n_waves) {
    const uint32_t wn = w + wstride;
    uint2 rn = make_uint2(0u, 0u);
    uint32_t usedn = 0;
    if (wn < b.n_waves) {
      rn = b.wave_lines[wn];
      usedn = b.wave_used[wn];
    }
    if (r.y != 0) {
      const size_t region = 2 * (size_t)r.x;
      // two groups of 64 entries in flight at once
      for (uint32_t j = threadIdx.x & 63; j < used; j += 128) {
        const bool h2 = j + 64 < used;
        const uint2 mt = b.list_meta[region + j];
        const float4 rc = b.list_rec[region + j];
        uint2 mt2 = make_uint2(0u, 0u);
        float4 rc2 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (h2) {
          mt2 = b.list_meta[region + j + 64];
          rc2 = b.list_rec[region + j + 64];
        }
        place_record(b, mt.x, mt.y, w, rc);
        if (h2) place_record(b, mt2.x, mt2.y, w, rc2);
      }
    }
    w = wn;
    r = rn;
    used = usedn;
  }
  const uint32_t stride = n_rec_blocks * kScatterBlock;
  const size_t obase = b.list_regions;
  for (uint32_t p = 0; p < kParts; ++p) {
    const uint32_t np = min(b.fill[p], b.part_cap);
    for (uint32_t k = blockIdx.x * kScatterBlock + threadIdx.x; k < np; k += stride) {
      const size_t o = (size_t)p * b.part_cap + k;
      const uint2 mt = b.list_meta[obase + o];
      if (mt.x != ~0u) place_record(b, mt.x, mt.y, b.ovf_wave[o], b.list_rec[obase + o]);
    }
  }
}

// Fold one run's status block into the context's accumulated one (first
// error, overflow bits, largest counts), so that every run queued before a
// sync is checked there: a later run that fits cannot hide an earlier one
// that did not. Called by one thread after all of the run's other kernels
// (stream order), usually by the first thread of its last kernel.
__device__ inline void merge_status(const Status* src, Status* acc) {
  if (!acc->err) acc->err = src->err;
  acc->overflow |= src->overflow;
  acc->n_records = max(acc->n_records, src->n_records);
  acc->n_events = max(acc->n_events, src->n_events);
  acc->n_lines = max(acc->n_lines, src->n_lines);
#pragma unroll
  for (int k = 0; k < 3; ++k) acc->pad[k] = max(acc->pad[k], src->pad[k]);
  acc->n_list = max(acc->n_list, src->n_list);
  acc->n_line_part = max(acc->n_line_part, src->n_line_part);
}

__global__ void k_status_merge(const Status* src, Status* acc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) merge_status(src, acc);
}

// Fill a layer with its background colour (Layer::new, src/scene.rs:483-501).
__global__ void k_fill_rgba(rzp::C4* out, int w, int h, size_t row_stride, rzp::C4 c) {
  const size_t n = (size_t)w * h;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t y = i / (size_t)w, x = i - y * (size_t)w;
    out[y * row_stride + x] = c;
  }
}

// ---------------------------------------------------------------------------
// Scans
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

constexpr int kScanItems = 8;  // per thread
constexpr int kScanTile = kBlock * kScanItems;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    uint32_t s = wsum[w];
    if (w < wid) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

// A whole exclusive scan done by one block (a short array riding along with
// another scan's second pass).
struct SmallScan {
  const uint32_t* in;
  uint32_t* out;     // [n + 1]
  uint32_t* cursor;  // copy of the offsets, or null
  uint32_t* total;   // or null
  uint32_t n;        // 0: none
};

// Pass 1: per-tile exclusive scan of `in` into `out`, block totals into sums.
__global__ __launch_bounds__(kBlock) void k_scan_tiles(const uint32_t* in, uint32_t* out, size_t n,
                                                        uint32_t* sums) {
  size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems], s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t pre = block_exclusive_scan(s, &tot);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = pre;
    pre += v[k];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Many block totals (more than kScanDirect): one block scans them in place
// and writes the grand total to out[n] and *total.
constexpr uint32_t kScanDirect = 4096;

__global__ __launch_bounds__(kBlock) void k_scan_sums(uint32_t* sums, size_t nb, uint32_t* out,
                                                       size_t n, uint32_t* total) {
  uint32_t carry = 0;
  for (size_t start = 0; start < nb; start += kBlock) {
    size_t i = start + threadIdx.x;
    uint32_t v = i < nb ? sums[i] : 0u;
    uint32_t tot;
    uint32_t pre = block_exclusive_scan(v, &tot);
    if (i < nb) sums[i] = carry + pre;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[n] = carry;
    if (total) *total = carry;
  }
}

// Sum of v over the block (every thread takes part).
__device__ inline uint32_t block_sum(uint32_t v) {
  uint32_t tot;
  (void)block_exclusive_scan(v, &tot);
  return tot;
}

// Exclusive scan of 64-bit values over the block; the total to *total.
__device__ inline unsigned long long block_exclusive_scan64(unsigned long long v,
                                                            unsigned long long* total) {
  __shared__ unsigned long long wsum64[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned long long x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long y = (unsigned long long)__shfl_up((long long)x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum64[wid] = x;
  __syncthreads();
  unsigned long long base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    const unsigned long long q = wsum64[w];
    if (w < wid) base += q;
    tot += q;
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

// Inclusive prefix along rows of a signed array (two's-complement sums, so
// the unsigned scan is exact): `rows` rows of `len` int32 values from `p`,
// then, if n_tail > 0, one more row of n_tail int64 values from `tail`.
struct RowPrefix {
  int32_t* p;
  int len;
  uint32_t rows;
  long long* tail;
  int n_tail;
};

__device__ void row_prefix_block(const RowPrefix& r, uint32_t row) {
  if (row == r.rows) {
    unsigned long long* v = (unsigned long long*)r.tail;
    const int n = r.n_tail;
    const int per = (n + kBlock - 1) / kBlock;
    const int i0 = (int)threadIdx.x * per, i1 = min(i0 + per, n);
    unsigned long long sum = 0;
    for (int i = i0; i < i1; ++i) sum += v[i];
    unsigned long long tot;
    unsigned long long run = block_exclusive_scan64(sum, &tot);
    for (int i = i0; i < i1; ++i) {
      run += v[i];
      v[i] = run;
    }
    return;
  }
  uint32_t* v = (uint32_t*)(r.p + (size_t)row * r.len);
  const int n = r.len;
  // each thread takes `per` consecutive values: one block scan of their sums
  const int per = (n + kBlock - 1) / kBlock;
  const int i0 = (int)threadIdx.x * per, i1 = min(i0 + per, n);
  uint32_t sum = 0;
  for (int i = i0; i < i1; ++i) sum += v[i];
  uint32_t tot;
  uint32_t run = block_exclusive_scan(sum, &tot);
  for (int i = i0; i < i1; ++i) {
    run += v[i];
    v[i] = run;
  }
}

// Pass 2: add block offsets; optionally copy to a cursor array. Unless the
// block totals were scanned in place (`scanned`), each block sums those
// before its own. The last block writes the grand total to out[n] and
// *total. The block after the scan's scans the short array `x` (if any), the
// blocks after that take the rows of `r`.
__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* out, size_t n, const uint32_t* sums,
                                                      uint32_t nb, bool scanned, uint32_t* cursor,
                                                      uint32_t* total, SmallScan x, RowPrefix r) {
  if (blockIdx.x >= nb) {
    uint32_t k = blockIdx.x - nb;
    if (x.n && k == 0) {
      uint32_t carry = 0;
      for (uint32_t start = 0; start < x.n; start += kBlock) {
        const uint32_t i = start + threadIdx.x;
        const uint32_t v = i < x.n ? x.in[i] : 0u;
        uint32_t tot;
        const uint32_t pre = block_exclusive_scan(v, &tot);
        if (i < x.n) {
          x.out[i] = carry + pre;
          if (x.cursor) x.cursor[i] = carry + pre;
        }
        carry += tot;
        __syncthreads();
      }
      if (threadIdx.x == 0) {
        x.out[x.n] = carry;
        if (x.total) *x.total = carry;
      }
      return;
    }
    if (x.n) --k;
    row_prefix_block(r, k);
    return;
  }
  uint32_t off;
  if (scanned) {
    off = sums[blockIdx.x];
  } else {
    uint32_t part = 0;
    for (uint32_t j = threadIdx.x; j < blockIdx.x; j += kBlock) part += sums[j];
    off = block_sum(part);
  }
  if (!scanned && blockIdx.x == nb - 1 && threadIdx.x == 0) {
    const uint32_t t = off + sums[blockIdx.x];
    out[n] = t;
    if (total) *total = t;
  }
  size_t base = (size_t)blockIdx.x * kScanTile;
  for (int k = threadIdx.x; k < kScanTile; k += kBlock) {
    size_t i = base + k;
    if (i < n) {
      uint32_t v = out[i] + off;
      out[i] = v;
      if (cursor) cursor[i] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Fine stage
// ---------------------------------------------------------------------------
constexpr float kFix = (float)(1 << rzg::kFixShift);

__device__ inline int fix(float v) { return __float2int_rn(v * kFix); }

// Inclusive prefix sum inside each 16-lane DPP row (row_shr:1,2,4,8).
__device__ inline int row_scan16(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
  return v;
}

// Add sign * V(yc) to the per-row base, V(yc)(row) = clamp(yc - row, 0, 1),
// stored as differences between consecutive rows (prefix-summed later). The
// difference of row 0 goes to the lane's own sum r0 (added to rowd[0] once
// per tile by flush_row0: an atomic on one address from many lanes would be
// serialized lane by lane).
__device__ inline void add_v(int* rowd, uint32_t& r0, float yc, float sign) {
  if (!(yc > 0.0f)) return;
  if (yc >= 16.0f) {
    r0 += (uint32_t)(int)(sign * kFix);
    return;
  }
  int k = (int)yc;
  float f = yc - (float)k;
  int one = (int)(sign * kFix), ff = fix(sign * f);
  if (k == 0) {
    r0 += (uint32_t)ff;
  } else {
    r0 += (uint32_t)one;
    atomicAdd(&rowd[k], ff - one);
  }
  atomicAdd(&rowd[k + 1], -ff);
}

// Add the lanes' row-0 sums to rowd[0] (two's-complement sums, exact in any
// order). Every lane of the wave takes part.
__device__ inline void flush_row0(int* rowd, uint32_t r0, int lane) {
  const uint32_t t = (uint32_t)__builtin_amdgcn_readlane(dpp_add_scan((int)r0), 63);
  if (lane == 0 && t) atomicAdd(&rowd[0], (int)t);
}

// The tile's accumulators are rows of kAccStride ints: with 17, the same
// column of different rows falls in different LDS banks (neighbouring lanes
// often deposit into neighbouring rows of one steep record).
constexpr int kAccStride = 17;
__device__ inline void dep(int* acc, int row, int col, float v) {
  if ((unsigned)col < 16u) atomicAdd(&acc[row * kAccStride + col], fix(v));
}

// signed_difference_line (src/rasterize.rs:395-469) for the cell records of
// one tile, in cell coordinates; deposits beyond column 15 belong to tiles to
// the right and are accounted for by their backdrops.
//
// A record crosses one to 16 pixel rows. The records are taken 64 at a time,
// one per lane (left-edge crossings, direction, slope and row range), and
// their (record, row) pairs are then dealt out to the lanes 64 at a time, so
// a record crossing many rows keeps many lanes busy rather than one. A row's
// x range is computed from the record's start and slope (x0 + dx/dy (y - y0)).

// A record prepared for its row steps, staged in LDS for the wave.
struct RowStage {
  float x0, y0, y1, dxdy, dir;
  int ys;   // first row
  uint32_t pre;  // first row step of the record among its round's steps
  int pad;
};

// Deposit the part of a prepared record within pixel row y.
__device__ inline void fine_row(int* acc, const RowStage& R, int y) {
  const float ylo = fmaxf((float)y, R.y0);
  const float x = R.x0 + R.dxdy * (ylo - R.y0);
  const float dy = fminf((float)(y + 1), R.y1) - ylo;
  const float d = R.dir * dy;
  const float x_next = x + R.dxdy * dy;
  const float xa = fminf(x, x_next), xb = fmaxf(x, x_next);
  const float xa_floor = fmaxf(floorf(xa), 0.0f);
  const int xai = (int)xa_floor;
  const float xb_ceil = ceilf(xb);
  const int xbi = (int)xb_ceil;
  if (xbi <= xai + 1) {
    const float xmf = 0.5f * (x + x_next) - xa_floor;
    dep(acc, y, xai, d * (1.0f - xmf));
    dep(acc, y, xai + 1, d * xmf);
  } else {
    const float s = 1.0f / (xb - xa);
    const float xaf = xa - xa_floor;
    const float xbf = xb - xb_ceil + 1.0f;
    const float a0 = 0.5f * s * (1.0f - xaf) * (1.0f - xaf);
    const float am = 0.5f * s * xbf * xbf;
    dep(acc, y, xai, d * a0);
    if (xbi == xai + 2) {
      dep(acc, y, xai + 1, d * (1.0f - a0 - am));
    } else {
      const float a1 = s * (1.5f - xaf);
      dep(acc, y, xai + 1, d * (a1 - a0));
      const int hi = xbi - 1 < 16 ? xbi - 1 : 16;
      for (int xi = xai + 2; xi < hi; ++xi) dep(acc, y, xi, d * s);
      const float a2 = a1 + (float)(xbi - xai - 3) * s;
      dep(acc, y, xbi - 1, d * (1.0f - a2 - am));
    }
    dep(acc, y, xbi, d * am);
  }
}

// Records [rb, re) of one tile into the wave's LDS accumulators (the row-0
// differences into the lane's row0, see add_v). `rs` and `marks` are the
// wave's 64-entry LDS scratch. Every lane of the wave takes part.
__device__ inline void tile_records(const float4* recs, uint32_t rb, uint32_t re, int* acc,
                                    int* rowd, uint32_t& row0, RowStage* rs, int* marks, int lane) {
  // each round's records are loaded while the previous round is deposited
  float4 nx = rb + (uint32_t)lane < re ? recs[rb + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t r0 = rb; r0 < re; r0 += 64) {
    const uint32_t i = r0 + (uint32_t)lane;
    int nrows = 0;
    RowStage R{};
    const float4 cur = nx;
    if (r0 + 64 + (uint32_t)lane < re) nx = recs[r0 + 64 + lane];
    if (i < re) {
      const float4 r = cur;
      float x0 = r.x, y0 = r.y, x1 = r.z, y1 = r.w;
      // records are clipped to the cell; anything else is skipped rather than
      // allowed to index outside the tile
      if (fabsf(x0) <= 16.0f && fabsf(x1) <= 16.0f && y0 >= 0.0f && y0 <= 16.0f && y1 >= 0.0f &&
          y1 <= 16.0f) {
        if (signbit(x0)) add_v(rowd, row0, y0, 1.0f);
        if (signbit(x1)) add_v(rowd, row0, y1, -1.0f);
        x0 = fabsf(x0);
        x1 = fabsf(x1);
        if (y0 != y1) {
          float dir = 1.0f;
          if (y0 > y1) {
            dir = -1.0f;
            float t = x0; x0 = x1; x1 = t;
            t = y0; y0 = y1; y1 = t;
          }
          const int ys = (int)y0;
          int ye = (int)ceilf(y1);
          if (ye > 16) ye = 16;
          nrows = ye - ys;
          R.x0 = x0;
          R.y0 = y0;
          R.y1 = y1;
          R.dxdy = (x1 - x0) / (y1 - y0);
          R.dir = dir;
          R.ys = ys;
        }
      }
    }
    uint32_t n_rows;
    R.pre = wave_excl_scan((uint32_t)nrows, lane, &n_rows);
    if (nrows > 0) rs[lane] = R;
    wave_lds_sync();
    int carry = -1;
    for (uint32_t k0 = 0; k0 < n_rows; k0 += 64) {
      const int o = task_owner(marks, R.pre, (uint32_t)nrows, k0, lane, &carry);
      const uint32_t k = k0 + (uint32_t)lane;
      if (k < n_rows) {
        const RowStage& Q = rs[o];
        fine_row(acc, Q, Q.ys + (int)(k - Q.pre));
      }
    }
    wave_lds_sync();  // the next round re-stages the records
  }
}

// The LDS of one tile's coverage (2^-20 fixed point, int32), and the int64
// per-pixel sums of a tile accumulated in chunks (lane-major: [k][lane]).
struct TileLDS {
  int acc[16 * kAccStride];
  int rowd[18];
  RowStage rs[64];
  int marks[64];
  long long w64[4 * 64];
};

// The int32 accumulation is exact while a tile's local winding (its records'
// deposits, their V terms and its band's events) stays inside int32. Per
// pixel, each record contributes at most 1 (its area in the row) plus 2 V
// terms of at most 1, each event at most 1, each term with rounding below
// 2^20 + 9: so |local| <= (3 records + events) (2^20 + 9) < 2^31 below 2000.
// Larger tiles are accumulated in chunks of at most 600 records and 190
// events, each chunk's int32 result folded into int64 per pixel.
constexpr uint32_t kChunkRecords = 600, kChunkEvents = 190;
__device__ inline bool tile_narrow(uint32_t n_records, uint32_t n_events) {
  return 3ull * n_records + n_events < 2000ull;
}

// One narrow chunk into LDS: records [rb, re), and those of the events
// [eb, ee) that lie left of tile column `tx` (in the path's grid); then the
// per-row base prefix. Every lane takes part.
__device__ inline void chunk_build(TileLDS& L, const float4* recs, uint32_t rb, uint32_t re,
                                   const int4* events, uint32_t eb, uint32_t ee, int tx, int lane) {
#pragma unroll
  for (int k = lane; k < 16 * kAccStride; k += 64) L.acc[k] = 0;
  if (lane < 18) L.rowd[lane] = 0;
  __syncthreads();
  uint32_t r0 = 0;
  tile_records(recs, rb, re, L.acc, L.rowd, r0, L.rs, L.marks, lane);
  for (uint32_t i = eb + lane; i < ee; i += 64) {
    const int4 ev = events[i];
    if (ev.y < tx) add_v(L.rowd, r0, __int_as_float(ev.z), __int_as_float(ev.w));
  }
  flush_row0(L.rowd, r0, lane);
  __syncthreads();
  // per-row base: prefix of the row differences (lanes 0..15 scan, in place)
  if (lane < 16) L.rowd[lane] = row_scan16(L.rowd[lane]);
  __syncthreads();
}

// The local winding of one tile (records [rb, re), the band's events [eb, ee))
// at each lane's four pixels (rows 4k + lane / 16, column lane % 16) into
// L.w64[k * 64 + lane], which tile_local reads back: the per-pixel sums stay in
// LDS, not in registers, while records are deposited. A narrow tile is one
// int32 chunk; a larger one is summed chunk by chunk. Every lane of the wave
// takes part (the row scans are DPP).
__device__ inline void tile_coverage(TileLDS& L, const float4* recs, uint32_t rb, uint32_t re,
                                     const int4* events, uint32_t eb, uint32_t ee, int tx,
                                     int lane) {
  const int col = lane & 15;
  const bool narrow = tile_narrow(re - rb, ee - eb);
  uint32_t r0 = rb, e0 = eb;
  bool first = true;
  do {
    const uint32_t r1 = narrow ? re : min(re, r0 + kChunkRecords);
    const uint32_t e1 = narrow ? ee : min(ee, e0 + kChunkEvents);
    chunk_build(L, recs, r0, r1, events, e0, e1, tx, lane);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int row = 4 * k + (lane >> 4);
      const long long v = (int32_t)(L.rowd[row] + row_scan16(L.acc[row * kAccStride + col]));
      L.w64[k * 64 + lane] = first ? v : L.w64[k * 64 + lane] + v;
    }
    first = false;
    r0 = r1;
    e0 = e1;
    // the next chunk clears the LDS; after the last, the barrier keeps the
    // caller's later uniform global loads scalar (the compiler takes an LDS
    // store not followed by a barrier as a possible clobber of them)
    __syncthreads();
  } while (r0 < re || e0 < ee);
}

// The local winding at the lane's pixel k after tile_coverage (each lane reads
// the slots it wrote: no barrier).
__device__ inline int64_t tile_local(const TileLDS& L, int k, int lane) { return L.w64[k * 64 + lane]; }

struct FineArgs {
  const uint32_t* tile_off;
  const float4* recs;
  const int32_t* bd;
  const uint32_t* band_ev_off;
  const int4* events;
  const long long* redge;  // [H] right-edge column per row (2^-20 fixed point)
  const Status* st;
  Status* acc;           // the context's accumulated status (merged by block 0)
  float* out;
  size_t row_stride;
  int W, H, tiles_x, tiles_y;
  int redge_col;         // floor(redge_x): gets redge_w * R; pixels right of it get R
  double redge_w;
};

// FillRule::alpha_from_winding (src/path.rs:32-46) on a 2^-20 fixed-point winding
// (the magnitudes converted to f32 are below 2^21, so the 32-bit conversion
// is exact and equals the 64-bit one)
template <int RULE>
__device__ inline float alpha_fix(int64_t w) {
  if (RULE == 0) {
    int64_t a = w < 0 ? -w : w;
    if (a >= ((int64_t)1 << rzg::kFixShift)) return 1.0f;
    float v = (float)(int32_t)a * (1.0f / kFix);
    return v < 1e-6f ? 0.0f : v;
  } else {
    // ((w + 1) rem_euclid 2 - 1).abs(): exact on the fixed-point value
    const int64_t one = (int64_t)1 << rzg::kFixShift;
    int64_t u = (w + one) & ((one << 1) - 1);
    int32_t d = (int32_t)(u - one);
    return (float)(d < 0 ? -d : d) * (1.0f / kFix);
  }
}

constexpr int kFineWave = 64;  // one wave per tile

// Per-pixel outputs of the fine stage.
// Rasterizer::mask: the coverage itself, every pixel of the image.
struct MaskOut {
  float* out;
  size_t row_stride;  // floats
  __device__ void put(int x, int y, float alpha) const { out[(size_t)y * row_stride + x] = alpha; }
};
// Rasterizer::fill (src/rasterize.rs:69-115): pixels that mask_iter yields
// (alpha >= 1e-6) get dst = dst.blend_over(paint.at(centre).with_alpha(alpha)).
struct FillOut {
  rzp::C4* img;
  size_t row_stride;  // pixels
  rzp::DPaint paint;
  __device__ void put(int x, int y, float alpha) const {
    if (!(alpha >= 1e-6f)) return;
    const rzp::C4 c = rzp::paint_at(paint, x, y);
    rzp::C4* d = img + (size_t)y * row_stride + x;
    *d = rzp::blend_over(*d, rzp::mul(c, alpha));
  }
};

// One 16x16 tile per wave (one workgroup per tile, so the hardware balances
// tiles of very different record counts): records and events into LDS, then
// four passes of 16-lane DPP row scans (one pixel row per 16 lanes) and
// coalesced stores.
template <int RULE, class Out>
__device__ inline void fine_tile(const FineArgs& a, const Out& o, TileLDS& L) {
  const int lane = threadIdx.x;
  const uint32_t tile = blockIdx.x;
  const int tx = (int)(tile % (uint32_t)a.tiles_x), ty = (int)(tile / (uint32_t)a.tiles_x);
  const uint32_t rb = a.tile_off[tile], re = a.tile_off[tile + 1];
  const uint32_t eb = a.band_ev_off[ty], ee = a.band_ev_off[ty + 1];
  // a tile with no records and no events of its band has the winding of its
  // backdrop (plus the right-edge column) everywhere: no LDS work
  const bool local = rb != re || eb != ee;
  const int64_t bd = (int64_t)a.bd[tile] << rzg::kFixShift;
  if (local) tile_coverage(L, a.recs, rb, re, a.events, eb, ee, tx, lane);
  const int col = lane & 15;
  const int x = tx * 16 + col;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int row = 4 * k + (lane >> 4);
    int64_t w = bd;
    if (local) w += tile_local(L, k, lane);
    const int y = ty * 16 + row;
    if (x < a.W && y < a.H) {
      if (x >= a.redge_col) {
        const int64_t r = a.redge[y];
        w += x == a.redge_col ? (int64_t)rint((double)r * a.redge_w) : r;
      }
      o.put(x, y, alpha_fix<RULE>(w));
    }
  }
}

// The status merge (one thread of block 0) comes after the tile's work: a
// global store ahead of the loads would keep the compiler from reading the
// tile's uniform data with scalar loads.
template <int RULE, class Out>
__global__ __launch_bounds__(kFineWave) void k_fine(FineArgs a, Out o) {
  __shared__ TileLDS L;
  if (!(a.st->err || a.st->overflow)) fine_tile<RULE>(a, o, L);
  if (blockIdx.x == 0 && threadIdx.x == 0) merge_status(a.st, a.acc);
}

}  // namespace rzk
