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
  uint32_t n_slots;
  uint32_t blocks_line, blocks_quad, blocks_cubic;  // block ranges per kind in `order`
};

struct Status {
  uint32_t err;        // first error (rzg::Status), 0 if none
  uint32_t overflow;   // bit0 records, bit1 events, bit2 lines, bit3 scene slices,
                       // bit4 a wave's long-line list outgrew its buffer
  uint32_t n_records;  // total records (written by the scan)
  uint32_t n_events;
  uint32_t n_lines;    // flattened lines (atomic counter of the flatten kernel)
  uint32_t pad[3];     // a batched scene's slice totals
  uint32_t n_long;     // longest long-line list of a wave, when one overflowed
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

// ---------------------------------------------------------------------------
// Flatten to lines for binning. Line slots take one thread each; every quad /
// cubic slot is split into 2^D0 virtual subtrees (the subtree at depth D0
// along path bits j), one thread each: the thread walks from the root along
// j; a leaf met above depth D0 is processed only by the leftmost j under it.
// Every leaf is produced exactly once, with the same arithmetic as the
// slot-wide traversal, and a curve with hundreds of lines is spread over many
// threads. The order of lines in the buffer does not matter for binning, so
// each wave reserves its output with one atomic.
// ---------------------------------------------------------------------------
struct LineBufs {
  double4* lines;
  uint32_t* line_path;  // batched scenes: path of each line
  uint32_t cap;
  uint32_t* count;  // total lines (device)
  Status* st;
};

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

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  return v;
}

// The subtree a thread of the flatten kernels handles. Returns the node to
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

// Count pass: lines per thread, and per wave their total (no atomics: the
// wave totals are scanned into the wave's output offsets).
template <int NP, bool MULTI>
__device__ void flat_count_item(const DPath& p, uint32_t slot, uint32_t j, int D0, bool active,
                                const Xform& x, double fl, const PathInfo* pi,
                                const uint32_t* slot_path, uint32_t* counts, uint32_t* wave_tot,
                                Status* st) {
  rzg::Node<NP> n;
  bool walk, leaf;
  int err;
  uint32_t pid;
  flat_subtree<NP, MULTI>(p, slot, j, D0, active, x, fl, pi, slot_path, &n, &walk, &leaf, &err, &pid);
  uint32_t cnt = leaf ? 1u : 0u;
  if (walk) {
    const int r = rzg::traverse(n, fl, [&](Pt, Pt) { ++cnt; });
    if (r != rzg::ST_OK) err = r;
  }
  if (err != rzg::ST_OK) {
    set_err(st, (uint32_t)err);
    cnt = 0;
  }
  const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
  counts[gid] = cnt;
  const uint32_t incl = wave_incl_scan(cnt);
  if ((threadIdx.x & 63) == 63) wave_tot[gid >> 6] = incl;
}

// Emit pass: the thread's lines at its wave's offset plus its prefix in the wave.
template <int NP, bool MULTI>
__device__ void flat_emit_item(const DPath& p, uint32_t slot, uint32_t j, int D0, bool active,
                               const Xform& x, double fl, const PathInfo* pi,
                               const uint32_t* slot_path, const uint32_t* counts,
                               const uint32_t* wave_off, const LineBufs& lb) {
  const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t cnt = counts[gid];
  const uint32_t incl = wave_incl_scan(cnt);
  if (!cnt) return;
  uint32_t o = wave_off[gid >> 6] + incl - cnt;
  if (o + cnt > lb.cap) {
    atomicOr(&lb.st->overflow, 4u);
    return;
  }
  rzg::Node<NP> n;
  bool walk, leaf;
  int err;
  uint32_t pid;
  flat_subtree<NP, MULTI>(p, slot, j, D0, active, x, fl, pi, slot_path, &n, &walk, &leaf, &err, &pid);
  if (leaf) {
    lb.lines[o] = make_double4(n.p[0].x, n.p[0].y, n.p[NP - 1].x, n.p[NP - 1].y);
    if (MULTI) lb.line_path[o] = pid;
    return;
  }
  rzg::traverse(n, fl, [&](Pt a, Pt c) {
    if (MULTI) lb.line_path[o] = pid;
    lb.lines[o++] = make_double4(a.x, a.y, c.x, c.y);
  });
}

// Blocks [0, bl) handle line slots, [bl, bl + bq) quad virtual items, the rest
// cubic virtual items. Every lane of a wave reaches the wave scans.
#define RZ_FLAT_DISPATCH(ITEM, ...)                                                        \
  const uint32_t blk = blockIdx.x, t = threadIdx.x;                                        \
  if (blk < bl) {                                                                          \
    const uint32_t slot = p.order[blk * kBlock + t];                                       \
    ITEM<2, MULTI>(p, slot, 0, 0, slot != ~0u, x, fl, pi, slot_path, __VA_ARGS__);         \
  } else if (blk < bl + bq) {                                                              \
    const uint32_t v = (blk - bl) * kBlock + t;                                            \
    const uint32_t slot = p.order[p.blocks_line * kBlock + (v >> dq)];                     \
    ITEM<3, MULTI>(p, slot, v & ((1u << dq) - 1u), dq, slot != ~0u, x, fl, pi, slot_path,  \
                   __VA_ARGS__);                                                           \
  } else {                                                                                 \
    const uint32_t v = (blk - bl - bq) * kBlock + t;                                       \
    const uint32_t slot = p.order[(p.blocks_line + p.blocks_quad) * kBlock + (v >> dc)];   \
    ITEM<4, MULTI>(p, slot, v & ((1u << dc) - 1u), dc, slot != ~0u, x, fl, pi, slot_path,  \
                   __VA_ARGS__);                                                           \
  }

template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_flat_count(DPath p, Xform x, double fl, uint32_t bl,
                                                        uint32_t bq, int dq, int dc,
                                                        const PathInfo* pi,
                                                        const uint32_t* slot_path,
                                                        uint32_t* counts, uint32_t* wave_tot,
                                                        Status* st) {
  if (st->overflow) return;  // a batched scene whose slices did not fit
  RZ_FLAT_DISPATCH(flat_count_item, counts, wave_tot, st)
}

template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_flat_emit(DPath p, Xform x, double fl, uint32_t bl,
                                                       uint32_t bq, int dq, int dc,
                                                       const PathInfo* pi, const uint32_t* slot_path,
                                                       const uint32_t* counts,
                                                       const uint32_t* wave_off, LineBufs lb) {
  if (lb.st->err || lb.st->overflow) return;
  RZ_FLAT_DISPATCH(flat_emit_item, counts, wave_off, lb)
}

// ---------------------------------------------------------------------------
// Binning of lines into cell records, one geometry pass per line:
//   k_bin_lines   counts records per tile, accumulates band-bottom crossings,
//                 right-edge rows and events per band, and keeps what it
//                 made: the records of a short line (at most two: 98% of
//                 flattened lines) in per-line slots, those of a longer line
//                 in a long-record list, events in an event list;
//   (scans of the tile and band counts)
//   k_bin_scatter moves the kept records into their tiles' slots;
//   k_bin_events  moves the listed events into their bands' slots.
// The lists are shared by all waves through kLongParts bump counters (a wave
// takes its entries with one atomic), so no single counter is contended.
// ---------------------------------------------------------------------------
constexpr uint32_t kLongParts = 16;

struct BinBufs {
  uint32_t* tile_count;   // counts; after the scan: cursors (start at the tile offsets)
  int32_t* bd_diff;       // [tiles] band backdrop differences
  uint32_t* band_ev;      // counts per band; after the scan: cursors
  float4* recs;
  uint32_t rec_cap;
  int4* events;           // (band, col, yrel bits, sign bits)
  uint32_t ev_cap;
  int32_t* redge;         // [H + 2] right-edge row differences
  uint2* ltile;           // [lines] tiles of a short line's records (~0u: none)
  float4* lrec;           // [2][lrec_stride] its first and second record
  uint32_t lrec_stride;
  uint32_t* long_tile;    // [kLongParts][long_cap] long lines' records: tile
  float4* long_rec;       //   ... and record
  uint32_t long_cap;
  uint32_t* fill;         // [kLongParts + 1] entries taken per long-record part, then events
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

// Exclusive prefix sum of v over the wave (every lane must take part); the
// wave total is returned in *total.
__device__ inline uint32_t wave_excl_scan(uint32_t v, int lane, uint32_t* total) {
  uint32_t s = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t u = (uint32_t)__shfl_up((int)s, d, 64);
    if (lane >= d) s += u;
  }
  *total = (uint32_t)__shfl((int)s, 63, 64);
  return s - v;
}

// Count pass, one thread per line (grid-stride, lanes of a wave on consecutive
// lines; the next line is loaded while this one is binned). A line with at
// most two records keeps them in registers; the wave then counts them per
// tile with one atomic per distinct tile (each round takes the tile of the
// first pending record and groups every pending record of that tile) and
// stores them in the line's slots. A longer line is binned a second time,
// counting and listing each record. Lines beyond the right clip only add to
// the right-edge column; two such lines on neighbouring lanes that continue
// each other drop their shared endpoint (its two terms cancel exactly).
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_bin_lines(const double4* lines,
                                                       const uint32_t* line_path,
                                                       const uint32_t* n_lines, rzg::BinGrid g0,
                                                       const PathInfo* pi, BinBufs b) {
  if (b.st->err || (b.st->overflow & 4u)) return;
  const uint32_t n = *n_lines;
  const uint32_t stride = gridDim.x * kBlock;
  const int lane = threadIdx.x & 63;
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t part = (blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) % kLongParts;
  const double4 zero = make_double4(0, 0, 0, 0);
  uint32_t first = blockIdx.x * kBlock + threadIdx.x;
  double4 Lnext = first < n ? lines[first] : zero;
  // the loop bound is uniform per wave so every lane reaches the wave operations
  for (uint32_t base = blockIdx.x * kBlock; base < n; base += stride) {
    const uint32_t i = base + threadIdx.x;
    const double4 L = Lnext;
    Lnext = i + stride < n ? lines[i + stride] : zero;
    int nrec = 0, t0 = -1, t1 = -1;
    rzg::Rec r0{}, r1{};  // separate registers: no dynamic indexing
    int nev = 0;
    int4 ev0{}, ev1{};
    uint32_t evb0 = 0, evb1 = 0;
    bool have = i < n;
    // this line's path: grid and slices of the per-tile / per-band / per-row arrays
    uint32_t pid = 0, tb = 0, bb = 0, rb0 = 0;
    rzg::BinGrid g = g0;
    if (MULTI && have) {
      pid = line_path[i];
      const PathInfo& P = pi[pid];
      g = P.g;
      tb = P.tile_base;
      bb = P.band_base;
      rb0 = P.redge_base;
    }
    const Pt p0 = rzg::mkpt(L.x, L.y), p1 = rzg::mkpt(L.z, L.w);
    if (have && !(isfinite(p0.x) && isfinite(p0.y) && isfinite(p1.x) && isfinite(p1.y))) {
      set_err(b.st, rzg::ST_NAN);
      have = false;
    }
    const bool proj = have && g.mode != rzg::EDGE_NONE && p0.x > g.wc && p1.x > g.wc;
    {
      const double nx = __shfl_down(L.x, 1, 64), ny = __shfl_down(L.y, 1, 64);
      const double px = __shfl_up(L.z, 1, 64), py = __shfl_up(L.w, 1, 64);
      const bool nproj = __shfl_down((int)proj, 1, 64) != 0;
      const bool pproj = __shfl_up((int)proj, 1, 64) != 0;
      const uint32_t npid = __shfl_down(pid, 1, 64), ppid = __shfl_up(pid, 1, 64);
      if (proj) {
        const bool skip_end = lane < 63 && nproj && npid == pid && nx == p1.x && ny == p1.y;
        const bool skip_start = lane > 0 && pproj && ppid == pid && px == p0.x && py == p0.y;
        auto add = [&](int k, int v) { atomicAdd(&b.redge[rb0 + k], v); };
        if (!skip_end) rzg::redge_endpoint(p1.y, 1, g, add);
        if (!skip_start) rzg::redge_endpoint(p0.y, -1, g, add);
      }
    }
    rzg::EffLine e{};
    if (have && !proj) {
      e = rzg::effective_line(p0, p1, g);
      auto collect = [&](int t, const rzg::Rec& r) {
        if (nrec == 0) {
          t0 = (int)tb + t;
          r0 = r;
        } else {
          t1 = (int)tb + t;  // only used when there are exactly two
          r1 = r;
        }
        ++nrec;
      };
      auto bd = [&](int band, int col, int dir) {
        atomicAdd(&b.bd_diff[tb + band * g.tiles_x + col], dir);
      };
      rzg::bin_piece(e.p[0], e.p[1], g, collect, bd);
      if (e.n == 2) rzg::bin_piece(e.p[1], e.p[2], g, collect, bd);
      rzg::Event ev;
      if (e.clip0 && rzg::make_event(e.p[0], -1.0f, g, &ev)) {
        ev0 = event_bits(ev);
        evb0 = bb + (uint32_t)ev.band;
        atomicAdd(&b.band_ev[evb0], 1u);
        nev = 1;
      }
      if (e.clip1 && rzg::make_event(e.n == 2 ? e.p[2] : e.p[1], 1.0f, g, &ev)) {
        if (nev == 0) {
          ev0 = event_bits(ev);
          evb0 = bb + (uint32_t)ev.band;
          atomicAdd(&b.band_ev[evb0], 1u);
        } else {
          ev1 = event_bits(ev);
          evb1 = bb + (uint32_t)ev.band;
          atomicAdd(&b.band_ev[evb1], 1u);
        }
        ++nev;
      }
    }
    const bool is_long = nrec > 2;
    // short lines: count per tile, wave-aggregated, and keep the records
    bool pend0 = !is_long && nrec >= 1, pend1 = !is_long && nrec == 2;
    while (true) {
      const uint64_t m0 = __ballot(pend0), m1 = __ballot(pend1);
      if ((m0 | m1) == 0) break;
      const bool use0 = m0 != 0;
      const int leader = __ffsll((unsigned long long)(use0 ? m0 : m1)) - 1;
      const int T = __builtin_amdgcn_readlane(use0 ? t0 : t1, leader);
      const bool mine0 = pend0 && t0 == T, mine1 = pend1 && t1 == T;
      const uint32_t total = (uint32_t)(__popcll(__ballot(mine0)) + __popcll(__ballot(mine1)));
      if (lane == leader) atomicAdd(&b.tile_count[T], total);
      pend0 = pend0 && !mine0;
      pend1 = pend1 && !mine1;
    }
    if (i < n) {
      const bool keep = !is_long && nrec > 0;
      b.ltile[i] = keep ? make_uint2((uint32_t)t0, nrec == 2 ? (uint32_t)t1 : ~0u)
                        : make_uint2(~0u, ~0u);
      if (keep) b.lrec[i] = make_float4(r0.x0, r0.y0, r0.x1, r0.y1);
      if (keep && nrec == 2) b.lrec[b.lrec_stride + i] = make_float4(r1.x0, r1.y0, r1.x1, r1.y1);
    }
    // long lines: take list entries for the wave, bin again, count and list
    uint32_t ltot;
    const uint32_t lpre = wave_excl_scan(is_long ? (uint32_t)nrec : 0u, lane, &ltot);
    if (ltot) {
      uint32_t lbase = 0;
      if (lane == 0) lbase = atomicAdd(&b.fill[part], ltot);
      lbase = (uint32_t)__shfl((int)lbase, 0, 64);
      // past the list's end the records are still counted (the record total
      // then sizes the re-run exactly)
      const bool fits = lbase + ltot <= b.long_cap;
      if (!fits && lane == 0) {
        atomicOr(&b.st->overflow, 16u);
        atomicMax(&b.st->n_long, lbase + ltot);
      }
      if (is_long) {
        size_t k = (size_t)part * b.long_cap + lbase + lpre;
        auto list = [&](int t, const rzg::Rec& r) {
          atomicAdd(&b.tile_count[tb + t], 1u);
          if (fits) {
            b.long_tile[k] = tb + (uint32_t)t;
            b.long_rec[k] = make_float4(r.x0, r.y0, r.x1, r.y1);
          }
          ++k;
        };
        auto nobd = [&](int, int, int) {};
        rzg::bin_piece(e.p[0], e.p[1], g, list, nobd);
        if (e.n == 2) rzg::bin_piece(e.p[1], e.p[2], g, list, nobd);
      }
    }
    // events: list them (at most two per line)
    const uint64_t e1 = __ballot(nev >= 1), e2 = __ballot(nev == 2);
    if (e1) {
      const uint32_t etot = (uint32_t)(__popcll(e1) + __popcll(e2));
      uint32_t ebase = 0;
      if (lane == 0) ebase = atomicAdd(&b.fill[kLongParts], etot);
      ebase = (uint32_t)__shfl((int)ebase, 0, 64);
      const uint32_t k = ebase + (uint32_t)(__popcll(e1 & below) + __popcll(e2 & below));
      if (ebase + etot > b.ev_cap) {
        if (lane == 0) atomicOr(&b.st->overflow, 2u);
      } else {
        if (nev >= 1) {
          b.ev_list[k] = ev0;
          b.ev_band[k] = evb0;
        }
        if (nev == 2) {
          b.ev_list[k + 1] = ev1;
          b.ev_band[k + 1] = evb1;
        }
      }
    }
  }
}

// Place the kept records into the slots the scan reserved: first the two
// slots of every line, then the long-record list, one entry per thread after
// the lines. The records of a wave that go to one tile take consecutive slots
// from a single atomic (ordered by lane, then record), and all of a wave's
// atomics are in flight together.
__global__ __launch_bounds__(kBlock) void k_bin_scatter(const uint32_t* n_lines, BinBufs b) {
  if (b.st->err || b.st->overflow) return;
  const uint32_t n = *n_lines;
  const uint32_t n_all = n + kLongParts * b.long_cap;
  const uint32_t stride = gridDim.x * kBlock;
  const int lane = threadIdx.x & 63;
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t base = blockIdx.x * kBlock; base < n_all; base += stride) {
    const uint32_t v = base + threadIdx.x;
    int nrec = 0, t0 = -1, t1 = -1;
    float4 r0 = make_float4(0, 0, 0, 0), r1 = r0;
    if (v < n) {
      const uint2 tt = b.ltile[v];
      t0 = (int)tt.x;
      t1 = (int)tt.y;
      nrec = tt.x == ~0u ? 0 : (tt.y == ~0u ? 1 : 2);
      if (nrec >= 1) r0 = b.lrec[v];
      if (nrec == 2) r1 = b.lrec[b.lrec_stride + v];
    } else if (v < n_all) {
      const uint32_t p = (v - n) / b.long_cap, j = (v - n) - p * b.long_cap;
      if (j < b.fill[p]) {
        t0 = (int)b.long_tile[v - n];
        r0 = b.long_rec[v - n];
        nrec = 1;
      }
    }
    bool pend0 = nrec >= 1, pend1 = nrec == 2;
    uint32_t tot0 = 0, tot1 = 0;      // group sizes this lane leads (for t0 / t1)
    int ldr0 = 0, ldr1 = 0;           // lane leading the group of record 0 / 1
    bool sel0 = false, sel1 = false;  // ... through its t1 atomic
    uint32_t off0 = 0, off1 = 0;
    while (true) {
      const uint64_t m0 = __ballot(pend0), m1 = __ballot(pend1);
      if ((m0 | m1) == 0) break;
      const bool use0 = m0 != 0;
      const int leader = __ffsll((unsigned long long)(use0 ? m0 : m1)) - 1;
      const int T = __builtin_amdgcn_readlane(use0 ? t0 : t1, leader);
      const bool mine0 = pend0 && t0 == T, mine1 = pend1 && t1 == T;
      const uint64_t g0 = __ballot(mine0), g1 = __ballot(mine1);
      const uint32_t total = (uint32_t)(__popcll(g0) + __popcll(g1));
      if (lane == leader) {
        if (use0) tot0 = total;
        else tot1 = total;
      }
      const uint32_t off = (uint32_t)(__popcll(g0 & below) + __popcll(g1 & below));
      if (mine0) {
        ldr0 = leader;
        sel0 = !use0;
        off0 = off;
        pend0 = false;
      }
      if (mine1) {
        ldr1 = leader;
        sel1 = !use0;
        off1 = off + (mine0 ? 1u : 0u);
        pend1 = false;
      }
    }
    uint32_t s0 = 0, s1 = 0;
    if (tot0) s0 = atomicAdd(&b.tile_count[t0], tot0);
    if (tot1) s1 = atomicAdd(&b.tile_count[t1], tot1);
    const uint32_t a0 = (uint32_t)__shfl((int)s0, ldr0, 64), b0 = (uint32_t)__shfl((int)s1, ldr0, 64);
    const uint32_t a1 = (uint32_t)__shfl((int)s0, ldr1, 64), b1 = (uint32_t)__shfl((int)s1, ldr1, 64);
    if (nrec >= 1) {
      const uint32_t slot = (sel0 ? b0 : a0) + off0;
      if (slot < b.rec_cap) b.recs[slot] = r0;
      else atomicOr(&b.st->overflow, 1u);
    }
    if (nrec == 2) {
      const uint32_t slot = (sel1 ? b1 : a1) + off1;
      if (slot < b.rec_cap) b.recs[slot] = r1;
      else atomicOr(&b.st->overflow, 1u);
    }
  }
}

// Place the events the count pass listed into their bands' slots.
__global__ __launch_bounds__(kBlock) void k_bin_events(BinBufs b) {
  if (b.st->err || b.st->overflow) return;
  const uint32_t n = b.fill[kLongParts];
  for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
    const uint32_t slot = atomicAdd(&b.band_ev[b.ev_band[k]], 1u);
    if (slot < b.ev_cap) b.events[slot] = b.ev_list[k];
    else atomicOr(&b.st->overflow, 2u);
  }
}

// A gap between the previous segment's end `a` and this segment's start `b`
// (arc endpoints are recomputed by sin/cos and miss their neighbours by
// ~1e-13): the chain has a free end at a (+V) and a free start at b (-V).
// Both events are made here; a pair that lands on the same band, column and
// f32 row offset contributes exactly zero and is dropped. Ends created by the
// right clip are reported by the clip itself. One thread per slot.
template <int PASS, bool MULTI>
__global__ __launch_bounds__(kBlock) void k_gap_events(DPath p, Xform x, rzg::BinGrid g0, BinBufs b,
                                                        const PathInfo* pi,
                                                        const uint32_t* slot_path) {
  if (b.st->err || b.st->overflow) return;
  const uint32_t slot = blockIdx.x * kBlock + threadIdx.x;
  if (slot >= p.n_slots || !(p.slot_flags[slot] & 1)) return;
  rzg::BinGrid g = g0;
  const double* m = x.m;
  uint32_t band_base = 0;
  if (MULTI) {
    const PathInfo& P = pi[slot_path[slot]];
    if (!P.live) return;
    g = P.g;
    m = P.m;
    band_base = P.band_base;
  }
  const uint32_t k = p.slot_item[slot];  // a segment (closing lines never start a gap)
  const double* qa = p.pts + 2 * (size_t)(p.pt_off[k] - 1);  // previous segment's last point
  const double* qb = p.pts + 2 * (size_t)p.pt_off[k];
  const Pt a = rzg::tr_apply(m, rzg::mkpt(qa[0], qa[1]));
  const Pt bb = rzg::tr_apply(m, rzg::mkpt(qb[0], qb[1]));
  const bool clip = g.mode != rzg::EDGE_NONE;
  const bool ea = !(clip && a.x > g.wc), eb = !(clip && bb.x > g.wc);
  rzg::Event va, vb;
  const Pt ap = rzg::mkpt(fmax(a.x, 0.0), a.y), bp = rzg::mkpt(fmax(bb.x, 0.0), bb.y);
  const bool ha = ea && rzg::make_event(ap, 1.0f, g, &va);
  const bool hb = eb && rzg::make_event(bp, -1.0f, g, &vb);
  if (ha && hb && va.band == vb.band && va.col == vb.col && va.yrel == vb.yrel) return;
  if (ha) put_event<PASS>(b, g, band_base, ap, 1.0f);
  if (hb) put_event<PASS>(b, g, band_base, bp, -1.0f);
}

// Fold one run's status block into a sticky one (first error, overflow bits,
// largest counts), for a sequence of runs that is checked once at the end.
__global__ void k_status_merge(const Status* src, Status* acc) {
  if (threadIdx.x != 0) return;
  if (!acc->err) acc->err = src->err;
  acc->overflow |= src->overflow;
  acc->n_records = max(acc->n_records, src->n_records);
  acc->n_events = max(acc->n_events, src->n_events);
  acc->n_lines = max(acc->n_lines, src->n_lines);
  acc->n_long = max(acc->n_long, src->n_long);
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

// Pass 2: one block scans the block totals (any count) and writes the grand
// total to out[n] and *total.
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

// Pass 3: add block offsets; optionally copy to a cursor array.
__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* out, size_t n, const uint32_t* sums,
                                                      uint32_t* cursor) {
  size_t base = (size_t)blockIdx.x * kScanTile;
  uint32_t off = sums[blockIdx.x];
  for (int k = threadIdx.x; k < kScanTile; k += kBlock) {
    size_t i = base + k;
    if (i < n) {
      uint32_t v = out[i] + off;
      out[i] = v;
      if (cursor) cursor[i] = v;
    }
  }
}

// Inclusive prefix of the band backdrop differences along each tile row, one
// block per band, and of the right-edge row differences in the last block
// (two's-complement sums, so the unsigned scan is exact).
__global__ __launch_bounds__(kBlock) void k_bd_prefix(int32_t* bd, int tiles_x, int32_t* redge,
                                                       int n_redge) {
  const bool edge = blockIdx.x == gridDim.x - 1;
  uint32_t* row = (uint32_t*)(edge ? redge : bd + (size_t)blockIdx.x * tiles_x);
  const int n = edge ? n_redge : tiles_x;
  uint32_t carry = 0;
  for (int start = 0; start < n; start += kBlock) {
    const int i = start + threadIdx.x;
    const uint32_t v = i < n ? row[i] : 0u;
    uint32_t tot;
    const uint32_t pre = block_exclusive_scan(v, &tot);
    if (i < n) row[i] = carry + pre + v;
    carry += tot;
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
// stored as differences between consecutive rows (prefix-summed later).
__device__ inline void add_v(int* rowd, float yc, float sign) {
  if (!(yc > 0.0f)) return;
  if (yc >= 16.0f) {
    atomicAdd(&rowd[0], (int)(sign * kFix));
    return;
  }
  int k = (int)yc;
  float f = yc - (float)k;
  int one = (int)(sign * kFix), ff = fix(sign * f);
  if (k == 0) {
    atomicAdd(&rowd[0], ff);
  } else {
    atomicAdd(&rowd[0], one);
    atomicAdd(&rowd[k], ff - one);
  }
  atomicAdd(&rowd[k + 1], -ff);
}

__device__ inline void dep(int* acc, int row, int col, float v) {
  if ((unsigned)col < 16u) atomicAdd(&acc[row * 16 + col], fix(v));
}

// signed_difference_line (src/rasterize.rs:395-469) for one cell record, in
// cell coordinates; deposits beyond column 15 belong to tiles to the right
// and are accounted for by their backdrops.
__device__ inline void fine_record(float4 r, int* acc, int* rowd) {
  float x0 = r.x, y0 = r.y, x1 = r.z, y1 = r.w;
  // records are clipped to the cell; anything else is skipped rather than
  // allowed to index outside the tile
  if (!(fabsf(x0) <= 16.0f && fabsf(x1) <= 16.0f && y0 >= 0.0f && y0 <= 16.0f && y1 >= 0.0f &&
        y1 <= 16.0f))
    return;
  if (signbit(x0)) add_v(rowd, y0, 1.0f);
  if (signbit(x1)) add_v(rowd, y1, -1.0f);
  x0 = fabsf(x0);
  x1 = fabsf(x1);
  if (y0 == y1) return;
  float dir = 1.0f;
  if (y0 > y1) {
    dir = -1.0f;
    float t = x0; x0 = x1; x1 = t;
    t = y0; y0 = y1; y1 = t;
  }
  const float dxdy = (x1 - x0) / (y1 - y0);
  int ys = (int)y0;
  int ye = (int)ceilf(y1);
  if (ye > 16) ye = 16;
  float x_next = x0;
  for (int y = ys; y < ye; ++y) {
    const float x = x_next;
    const float dy = fminf((float)(y + 1), y1) - fmaxf((float)y, y0);
    const float d = dir * dy;
    x_next = x + dxdy * dy;
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
}

struct FineArgs {
  const uint32_t* tile_off;
  const float4* recs;
  const int32_t* bd;
  const uint32_t* band_ev_off;
  const int4* events;
  const int32_t* redge;  // [H] right-edge column per row (2^-20 fixed point)
  const Status* st;
  float* out;
  size_t row_stride;
  int W, H, tiles_x, tiles_y;
  int redge_col;         // floor(redge_x): gets redge_w * R; pixels right of it get R
  double redge_w;
};

// FillRule::alpha_from_winding (src/path.rs:32-46) on a 2^-20 fixed-point winding
template <int RULE>
__device__ inline float alpha_fix(int64_t w) {
  if (RULE == 0) {
    int64_t a = w < 0 ? -w : w;
    if (a >= ((int64_t)1 << rzg::kFixShift)) return 1.0f;
    float v = (float)a * (1.0f / kFix);
    return v < 1e-6f ? 0.0f : v;
  } else {
    // ((w + 1) rem_euclid 2 - 1).abs(): exact on the fixed-point value
    const int64_t one = (int64_t)1 << rzg::kFixShift;
    int64_t u = (w + one) & ((one << 1) - 1);
    int64_t d = u - one;
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
__global__ __launch_bounds__(kFineWave) void k_fine(FineArgs a, Out o) {
  if (a.st->err || a.st->overflow) return;
  __shared__ int acc[256];
  __shared__ int rowd[18];
  const int lane = threadIdx.x;
  const uint32_t tile = blockIdx.x;
  const int tx = (int)(tile % (uint32_t)a.tiles_x), ty = (int)(tile / (uint32_t)a.tiles_x);
#pragma unroll
  for (int k = 0; k < 4; ++k) acc[lane + 64 * k] = 0;
  if (lane < 18) rowd[lane] = 0;
  __syncthreads();
  const uint32_t rb = a.tile_off[tile], re = a.tile_off[tile + 1];
  for (uint32_t i = rb + lane; i < re; i += kFineWave) fine_record(a.recs[i], acc, rowd);
  const uint32_t eb = a.band_ev_off[ty], ee = a.band_ev_off[ty + 1];
  for (uint32_t i = eb + lane; i < ee; i += kFineWave) {
    const int4 ev = a.events[i];
    if (ev.y < tx) add_v(rowd, __int_as_float(ev.z), __int_as_float(ev.w));
  }
  __syncthreads();
  // per-row base: prefix of the row differences (lanes 0..15 scan, in place)
  if (lane < 16) rowd[lane] = row_scan16(rowd[lane]);
  __syncthreads();
  const int64_t bd = (int64_t)a.bd[tile] << rzg::kFixShift;
  const int col = lane & 15;
  const int x = tx * 16 + col;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int row = 4 * k + (lane >> 4);
    const int v = row_scan16(acc[row * 16 + col]);
    int64_t w = bd + (int64_t)(int32_t)(rowd[row] + v);
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

}  // namespace rzk
