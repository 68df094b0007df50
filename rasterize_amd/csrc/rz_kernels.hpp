// HIP kernels of the mask pipeline for gfx950:
//   flatten  -- per-segment subdivision tree walk (f64, bit-exact)
//   bin      -- flatten fused with cell binning: per-tile records, band
//               backdrops, free-endpoint events (count pass, scatter pass)
//   fine     -- per-tile LDS accumulation of signed area/cover, DPP row scan,
//               fill rule, coalesced f32 stores
#pragma once
#include <hip/hip_runtime.h>
#include "rz_geom.hpp"

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
  uint32_t overflow;   // record or event buffer overflow
  uint32_t n_records;  // total records (written by the scan)
  uint32_t n_events;
};

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
// Binning (flatten fused with cell binning). PASS 0 counts records per tile,
// accumulates band-bottom crossings and counts events per band; PASS 1 writes
// records and events into the slots reserved by the scan.
// ---------------------------------------------------------------------------
struct BinBufs {
  uint32_t* tile_count;   // PASS 0: counts; PASS 1: cursor (starts at tile offsets)
  int32_t* bd_diff;       // [tiles] band backdrop differences
  uint32_t* band_ev;      // PASS 0: counts per band; PASS 1: cursor per band
  float4* recs;
  uint32_t rec_cap;
  int4* events;           // (band, col, yrel bits, sign bits)
  uint32_t ev_cap;
  Status* st;
};

template <int PASS>
struct BinSink {
  const rzg::BinGrid& g;
  const BinBufs& b;
  int run_tile;
  uint32_t run_n;
  __device__ BinSink(const rzg::BinGrid& g_, const BinBufs& b_) : g(g_), b(b_), run_tile(-1), run_n(0) {}
  __device__ void flush() {
    if (PASS == 0 && run_tile >= 0 && run_n) atomicAdd(&b.tile_count[run_tile], run_n);
    run_tile = -1;
    run_n = 0;
  }
  __device__ void rec(int tile, const rzg::Rec& r) {
    if (PASS == 0) {
      if (tile != run_tile) {
        flush();
        run_tile = tile;
      }
      ++run_n;
    } else {
      uint32_t slot = atomicAdd(&b.tile_count[tile], 1u);
      if (slot < b.rec_cap) b.recs[slot] = make_float4(r.x0, r.y0, r.x1, r.y1);
      else atomicOr(&b.st->overflow, 1u);
    }
  }
  __device__ void bd(int band, int col, int dir) {
    if (PASS == 0) atomicAdd(&b.bd_diff[band * g.tiles_x + col], dir);
  }
  __device__ void event(Pt p, float sign) {
    rzg::Event ev;
    if (!rzg::make_event(p, sign, g, &ev)) return;
    if (PASS == 0) {
      atomicAdd(&b.band_ev[ev.band], 1u);
    } else {
      uint32_t slot = atomicAdd(&b.band_ev[ev.band], 1u);
      if (slot < b.ev_cap)
        b.events[slot] = make_int4(ev.band, ev.col, __float_as_int(ev.yrel), __float_as_int(ev.sign));
      else atomicOr(&b.st->overflow, 2u);
    }
  }
  __device__ void line(Pt p0, Pt p1) {
    // infinite coordinates (the NaN case is caught by the traversal) would
    // produce NaN cell coordinates; report them like NaN
    if (!(isfinite(p0.x) && isfinite(p0.y) && isfinite(p1.x) && isfinite(p1.y))) {
      set_err(b.st, rzg::ST_NAN);
      return;
    }
    rzg::EffLine e = rzg::effective_line(p0, p1, g);
    for (int k = 0; k < e.n; ++k)
      rzg::bin_piece(e.p[k], e.p[k + 1], g, [&](int t, const rzg::Rec& r) { rec(t, r); },
                     [&](int band, int col, int dir) { bd(band, col, dir); });
    if (e.clip0) event(e.p[0], -1.0f);
    if (e.clip1) event(e.p[e.n], 1.0f);
  }
};

template <int NP, int PASS>
__device__ void bin_slot(const DPath& p, uint32_t slot, const Xform& x, double fl,
                         const rzg::BinGrid& g, const BinBufs& b) {
  rzg::Node<NP> root = load_node<NP>(p, slot, x.m);
  BinSink<PASS> sink(g, b);
  int r = rzg::traverse(root, fl, [&](Pt a, Pt c) { sink.line(a, c); });
  sink.flush();
  if (PASS == 0 && r != rzg::ST_OK) set_err(b.st, (uint32_t)r);
  // gaps between consecutive segments: the segment's own start/end is a free
  // chain endpoint unless the right clip already made it one
  uint8_t f = p.slot_flags[slot];
  if (f) {
    const double wc = g.wc;
    const bool clip = g.mode != rzg::EDGE_NONE;
    Pt s = root.p[0], e = root.p[NP - 1];
    if ((f & 1) && !(clip && s.x > wc)) sink.event(rzg::mkpt(fmax(s.x, 0.0), s.y), -1.0f);
    if ((f & 2) && !(clip && e.x > wc)) sink.event(rzg::mkpt(fmax(e.x, 0.0), e.y), 1.0f);
  }
}

template <int PASS>
__global__ __launch_bounds__(kBlock) void k_bin(DPath p, Xform x, double fl, rzg::BinGrid g,
                                                 BinBufs b) {
  if (PASS == 1 && b.st->err) return;
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  uint32_t slot = p.order[i];
  if (slot == ~0u) return;
  switch (block_kind(p, blockIdx.x)) {
    case rzg::KIND_LINE: bin_slot<2, PASS>(p, slot, x, fl, g, b); break;
    case rzg::KIND_QUAD: bin_slot<3, PASS>(p, slot, x, fl, g, b); break;
    default: bin_slot<4, PASS>(p, slot, x, fl, g, b); break;
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

// Inclusive prefix of the band backdrop differences along each tile row.
__global__ void k_bd_prefix(int32_t* bd, int tiles_x, int tiles_y) {
  int band = blockIdx.x * blockDim.x + threadIdx.x;
  if (band >= tiles_y) return;
  int32_t* row = bd + (size_t)band * tiles_x;
  int32_t acc = 0;
  for (int i = 0; i < tiles_x; ++i) {
    acc += row[i];
    row[i] = acc;
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
  const Status* st;
  float* out;
  size_t row_stride;
  int W, H, tiles_x, tiles_y;
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

template <int RULE>
__global__ __launch_bounds__(kBlock) void k_fine_mask(FineArgs a) {
  if (a.st->err || a.st->overflow) return;
  const int tile = blockIdx.x;
  const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
  __shared__ int acc[256];
  __shared__ int rowd[18];
  const int tid = threadIdx.x;
  acc[tid] = 0;
  if (tid < 18) rowd[tid] = 0;
  __syncthreads();
  const uint32_t rb = a.tile_off[tile], re = a.tile_off[tile + 1];
  for (uint32_t i = rb + tid; i < re; i += kBlock) fine_record(a.recs[i], acc, rowd);
  const uint32_t eb = a.band_ev_off[ty], ee = a.band_ev_off[ty + 1];
  for (uint32_t i = eb + tid; i < ee; i += kBlock) {
    int4 ev = a.events[i];
    if (ev.y < tx) add_v(rowd, __int_as_float(ev.z), __int_as_float(ev.w));
  }
  __syncthreads();
  const int row = tid >> 4, col = tid & 15;
  int v = row_scan16(acc[tid]);
  int base = 0;
  for (int k = 0; k <= row; ++k) base += rowd[k];
  const int64_t w = ((int64_t)a.bd[tile] << rzg::kFixShift) + (int64_t)(int32_t)(base + v);
  const int x = tx * 16 + col, y = ty * 16 + row;
  if (x < a.W && y < a.H) a.out[(size_t)y * a.row_stride + x] = alpha_fix<RULE>(w);
}

}  // namespace rzk
