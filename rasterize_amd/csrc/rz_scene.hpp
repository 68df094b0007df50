// HIP kernels of the batched scene render: Scene::group(fills).render(...)
// (src/scene.rs:186-199, 268-357, 384-459) for thousands of paths in one
// pass. Per render:
//   k_scene_bbox    -- Path::bbox(tr) of every path (a lane per segment)
//   k_scene_setup   -- Pipeline::build_rec's view intersection and
//                      render_rec's sub-view and alignment per path, its
//                      binning grid on the canvas tile grid, its paint
//   k_scene_offsets -- slices of the per-tile / per-band / per-row arrays
//   (flatten and binning kernels of the mask pipeline, in their batched form)
//   k_scene_bd      -- band backdrop and right-edge prefixes per path
//   k_scene_rows    -- per canvas tile row, the paths crossing it, in order
//   k_scene_fine    -- per canvas tile: every path over it in order, its
//                      coverage, paint and source-over; one store per pixel
#pragma once
#include "rz_host.hpp"
#include "rz_kernels.hpp"

namespace rzk {

// Static per-path description, uploaded once with the scene.
struct ScenePath {
  uint32_t seg_begin, seg_end;  // segments in the combined path
  int rule;
  int has_tr;                   // a Transform node above the fill: draw under render tr * tr
  double tr[6];
  int has_bbox_i;               // Path::bbox(identity) exists
  rzh::BBox bbox_i;             // for BoundingBox paints
  int units;
  double paint_tr[6];
  rzp::DPaint paint;            // stops already resolved (device pointers)
};

struct SceneView {
  double vx0, vy0, vx1, vy1;  // view
  int lx, ly, lw, lh;          // layer origin and size (Layer::new, src/scene.rs:483-501)
  double tr[6];
  int mode;                    // edge mode of the rasterizer's mask_iter
};

// Path::bbox(tr) of every path as ordered keys, so that the union is an
// integer max: key[4i + 0, 1] = ~okey(min x, min y), key[4i + 2, 3] =
// okey(max x, max y), all starting at zero. The union of segment boxes does
// not depend on the fold order (a quad or cubic whose inner control points
// already lie in the running box has its curve inside it), so each lane takes
// one segment (closing lines lie inside the segments' box) and each path of
// the wave gets one atomic per key. NaN coordinates drop out as in fmin/fmax.
__device__ inline unsigned long long okey(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double okey_value(unsigned long long k) {
  return __longlong_as_double((long long)((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k));
}

// The transform a fill is drawn under: the render's, times the fill's own
// (Pipeline::build_rec's tr * child_tr, src/scene.rs:341-344) if it has one.
__device__ inline rzh::Tr fill_tr(const SceneView& v, const ScenePath& S) {
  return S.has_tr ? rzh::tr_mul(rzh::tr_of(v.tr), rzh::tr_of(S.tr)) : rzh::tr_of(v.tr);
}

__global__ __launch_bounds__(64) void k_scene_bbox(DPath p, const uint32_t* slot_path, SceneView v,
                                                    const ScenePath* sp, unsigned long long* keys) {
  const uint32_t slot = blockIdx.x * 64 + threadIdx.x;
  const int lane = threadIdx.x;
  bool have = false;
  uint32_t pid = 0;
  double mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
  if (slot < p.n_slots) {
    const uint32_t k = p.slot_item[slot];
    if (!(k & 0x80000000u)) {
      have = true;
      pid = slot_path[slot];
      const int kind = p.kinds[k];
      const double* q = p.pts + 2 * (size_t)p.pt_off[k];
      const rzh::Tr m = fill_tr(v, sp[pid]);
      Pt c[4];
      for (int j = 0; j < kind + 2; ++j) c[j] = rzg::tr_apply(m.m, rzg::mkpt(q[2 * j], q[2 * j + 1]));
      const rzh::BBox b = rzh::segment_bbox(kind, c, nullptr);
      mnx = fmin(mnx, b.mn.x);
      mny = fmin(mny, b.mn.y);
      mxx = fmax(mxx, b.mx.x);
      mxy = fmax(mxy, b.mx.y);
    }
  }
  unsigned long long k0 = ~okey(mnx), k1 = ~okey(mny), k2 = okey(mxx), k3 = okey(mxy);
  uint64_t todo = __ballot(have);
  while (todo) {
    const int first = __ffsll((unsigned long long)todo) - 1;
    const uint32_t q = (uint32_t)__builtin_amdgcn_readlane((int)pid, first);
    const bool mine = have && pid == q;
    todo &= ~__ballot(mine);
    unsigned long long a0 = mine ? k0 : 0ull, a1 = mine ? k1 : 0ull;
    unsigned long long a2 = mine ? k2 : 0ull, a3 = mine ? k3 : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      a0 = max(a0, (unsigned long long)__shfl_xor((long long)a0, o, 64));
      a1 = max(a1, (unsigned long long)__shfl_xor((long long)a1, o, 64));
      a2 = max(a2, (unsigned long long)__shfl_xor((long long)a2, o, 64));
      a3 = max(a3, (unsigned long long)__shfl_xor((long long)a3, o, 64));
    }
    if (lane == first) {
      atomicMax(&keys[4 * (size_t)q + 0], a0);
      atomicMax(&keys[4 * (size_t)q + 1], a1);
      atomicMax(&keys[4 * (size_t)q + 2], a2);
      atomicMax(&keys[4 * (size_t)q + 3], a3);
    }
  }
}

// Per path: bbox ∩ view (Pipeline::build_rec, src/scene.rs:276-293), the
// sub-view rows/cols [floor(min) - layer, ceil(max) - layer + 1) clamped to
// the layer (src/scene.rs:413-423, src/image.rs:588-605), the alignment
// translate(-floor(min)) (:425-428) and the paint's pixel transform for the
// aligned fill. Writes the counts that k_scene_offsets turns into slices.
__global__ __launch_bounds__(kBlock) void k_scene_setup(const ScenePath* sp, uint32_t n,
                                                         SceneView v, const unsigned long long* keys,
                                                         PathInfo* pi, uint32_t* counts,
                                                         int4* rects) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const ScenePath& S = sp[i];
  PathInfo P;
  P.live = 0;
  P.rule = S.rule;
  uint32_t nt = 0, nb = 0, nr = 0;
  int4 rect = make_int4(0, 0, 0, 0);
  // the path's box (k_scene_bbox); a key still zero saw only NaN coordinates
  const unsigned long long* kk = keys + 4 * (size_t)i;
  double4 b;
  b.x = kk[0] ? okey_value(~kk[0]) : INFINITY;
  b.y = kk[1] ? okey_value(~kk[1]) : INFINITY;
  b.z = kk[2] ? okey_value(kk[2]) : -INFINITY;
  b.w = kk[3] ? okey_value(kk[3]) : -INFINITY;
  const bool has = S.seg_end > S.seg_begin;
  // range_intersect (src/geometry.rs:677-688) of view and bbox, both axes
  const bool hit = has && !(v.vx0 > b.z || b.x > v.vx1) && !(v.vy0 > b.w || b.y > v.vy1);
  if (hit) {
    const double bx0 = fmax(v.vx0, b.x), bx1 = fmin(v.vx1, b.z);
    const double by0 = fmax(v.vy0, b.y), by1 = fmin(v.vy1, b.w);
    const double fx = floor(bx0), fy = floor(by0);
    int col_min = (int)fx - v.lx, col_max = (int)ceil(bx1) - v.lx + 1;
    int row_min = (int)fy - v.ly, row_max = (int)ceil(by1) - v.ly + 1;
    col_min = min(max(col_min, 0), v.lw);
    col_max = min(max(col_max, col_min), v.lw);
    row_min = min(max(row_min, 0), v.lh);
    row_max = min(max(row_max, row_min), v.lh);
    const int W = col_max - col_min, H = row_max - row_min;
    rzh::Tr align;
    const double av[6] = {1.0, 0.0, -fx, 0.0, 1.0, -fy};
    for (int k = 0; k < 6; ++k) align.m[k] = av[k];
    const rzh::Tr m = rzh::tr_mul(align, fill_tr(v, S));
    for (int k = 0; k < 6; ++k) P.m[k] = m.m[k];
    P.paint = S.paint;
    bool draw = W > 0 && H > 0;
    if (draw && S.paint.kind != rzp::PAINT_SOLID)
      draw = rzh::paint_pixel_transform(S.units, S.paint_tr, m, S.has_bbox_i ? &S.bbox_i : nullptr,
                                        P.paint.ptr);
    if (draw) {
      P.live = 1;
      P.ox = col_min;
      P.oy = row_min;
      P.g.mode = v.mode;
      P.g.W = W;
      P.g.H = H;
      P.g.wc = v.mode == rzg::EDGE_MASK ? (double)W - 1.0 : (double)W;
      P.g.phx = col_min & 15;
      P.g.phy = row_min & 15;
      P.g.tiles_x = (W + P.g.phx + 15) >> 4;
      P.g.tiles_y = (H + P.g.phy + 15) >> 4;
      P.tx0 = col_min >> 4;
      P.ty0 = row_min >> 4;
      if (v.mode == rzg::EDGE_NONE) {
        P.redge_col = INT32_MAX;
        P.redge_w = 0.0;
      } else {
        const double X = rzg::redge_x(P.g), fxr = floor(X);
        P.redge_col = (int)fxr;
        P.redge_w = 1.0 - (X - fxr);
      }
      nt = (uint32_t)(P.g.tiles_x * P.g.tiles_y);
      nb = (uint32_t)P.g.tiles_y;
      nr = (uint32_t)H + 2u;
      rect = make_int4(P.tx0, P.ty0, P.tx0 + P.g.tiles_x, P.ty0 + P.g.tiles_y);
    }
  }
  pi[i] = P;
  counts[i] = nt;
  counts[n + i] = nb;
  counts[2 * n + i] = nr;
  rects[i] = rect;
}

// Exclusive prefix of the three per-path counts (one block), totals into the
// status block; a total beyond its buffer flags overflow (the render is then
// re-run with grown buffers).
__global__ __launch_bounds__(kBlock) void k_scene_offsets(uint32_t n, const uint32_t* counts,
                                                           PathInfo* pi, Status* st, uint32_t cap_t,
                                                           uint32_t cap_b, uint32_t cap_r) {
  // each thread takes a run of consecutive paths: one block scan of the run
  // totals gives each run's first offsets
  const uint32_t per = (n + kBlock - 1) / kBlock;
  const uint32_t i0 = min(threadIdx.x * per, n), i1 = min(i0 + per, n);
  uint32_t run[3] = {0, 0, 0}, carry[3];
  for (uint32_t i = i0; i < i1; ++i)
    for (int c = 0; c < 3; ++c) run[c] += counts[c * n + i];
  for (int c = 0; c < 3; ++c) {
    const uint32_t pre = block_exclusive_scan(run[c], &carry[c]);
    run[c] = pre;
  }
  for (uint32_t i = i0; i < i1; ++i) {
    pi[i].tile_base = run[0];
    pi[i].band_base = run[1];
    pi[i].redge_base = run[2];
    for (int c = 0; c < 3; ++c) run[c] += counts[c * n + i];
  }
  if (threadIdx.x == 0) {
    st->pad[0] = carry[0];
    st->pad[1] = carry[1];
    st->pad[2] = carry[2];
    if (carry[0] > cap_t || carry[1] > cap_b || carry[2] > cap_r) atomicOr(&st->overflow, 8u);
  }
}

// Band-backdrop prefix along each tile row and right-edge prefix over the
// rows (inclusive, two's-complement sums), one wave per path: lane r walks
// tile row r; for the right-edge column each lane takes a run of consecutive
// rows, and one wave scan of the run totals gives each run's carry.
__global__ __launch_bounds__(64) void k_scene_bd(const PathInfo* pi, int32_t* bd, long long* redge,
                                                  const Status* st) {
  if (st->err || st->overflow) return;
  const PathInfo& P = pi[blockIdx.x];
  if (!P.live) return;
  const int lane = threadIdx.x;
  const int tiles_x = P.g.tiles_x;
  for (int band = lane; band < P.g.tiles_y; band += 64) {
    uint32_t* row = (uint32_t*)(bd + P.tile_base + (size_t)band * tiles_x);
    uint32_t carry = 0;
#pragma unroll 8
    for (int i = 0; i < tiles_x; ++i) {
      carry += row[i];
      row[i] = carry;
    }
  }
  unsigned long long* e = (unsigned long long*)(redge + P.redge_base);
  const int n = P.g.H + 2;
  const int per = (n + 63) / 64;
  const int i0 = lane * per, i1 = min(i0 + per, n);
  unsigned long long sum = 0;
  for (int i = i0; i < i1; ++i) sum += e[i];
  unsigned long long incl = sum;  // inclusive wave scan, then exclusive
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long y = (unsigned long long)__shfl_up((long long)incl, o, 64);
    if (lane >= o) incl += y;
  }
  unsigned long long run = incl - sum;
  for (int i = i0; i < i1; ++i) {
    run += e[i];
    e[i] = run;
  }
}

// For canvas tile row `ty`, the live paths whose tiles cross it, in path
// order: rows[ty * n + k], count in row_count[ty]. One wave per row.
__global__ __launch_bounds__(64) void k_scene_rows(const int4* rects, uint32_t n, uint32_t* rows,
                                                    uint32_t* row_count, const Status* st) {
  if (st->err || st->overflow) return;
  const int ty = blockIdx.x;
  const int lane = threadIdx.x;
  const uint64_t below = (1ull << lane) - 1ull;
  constexpr int kBatch = 8;  // rows of 64 paths whose rects are loaded together
  uint32_t count = 0;
  for (uint32_t start = 0; start < n; start += 64 * kBatch) {
    int4 r[kBatch];
#pragma unroll
    for (int b = 0; b < kBatch; ++b) {
      const uint32_t i = start + 64 * b + lane;
      r[b] = i < n ? rects[i] : make_int4(0, 0, 0, 0);
    }
#pragma unroll
    for (int b = 0; b < kBatch; ++b) {
      const uint32_t i = start + 64 * b + lane;
      const bool hit = i < n && r[b].y <= ty && ty < r[b].w;
      const uint64_t m = __ballot(hit);
      if (hit) rows[(size_t)ty * n + count + (uint32_t)__popcll(m & below)] = i;
      count += (uint32_t)__popcll(m);
    }
  }
  if (lane == 0) row_count[ty] = count;
}

struct SceneFineArgs {
  const PathInfo* pi;
  const int4* rects;
  const uint32_t* rows;
  const uint32_t* row_count;
  uint32_t n_paths;
  const uint32_t* tile_off;
  const float4* recs;
  const int32_t* bd;
  const uint32_t* band_ev_off;
  const int4* events;
  const long long* redge;
  const Status* st;
  Status* acc;        // accumulated status (merged by block 0), or null
  rzp::C4 bg;
  int load_dst;       // start from the layer's pixels (a layer already drawn into) instead of bg
  rzp::C4* out;       // layer image
  size_t row_stride;  // pixels
  int lw, lh;         // layer size
  int tiles_x;        // canvas tiles per row
  const uint32_t* order;  // queued form: canvas tiles, costliest first
  uint32_t n_tiles;
};

// One canvas tile per wave. Each lane owns 4 pixels (rows 4k + lane/16,
// column lane % 16) and keeps their colours in registers from the background
// to the final store. Paths crossing the tile are taken in path order; each
// one's coverage is accumulated like k_fine's (records and events into LDS,
// DPP row scans) in its own sub-view frame, then its paint is composited
// source-over where mask_iter would yield a pixel (alpha >= 1e-6).
//
// The paths of the tile row are looked at 64 at a time, one per lane, which
// also loads the path's tile slot, record and event ranges and backdrop, so
// the path-by-path loop starts from values already in registers. A path with
// no records or events in the tile, a zero backdrop and no pixel at or right
// of its right-edge column has zero winding on the whole tile (alpha 0
// everywhere) and is skipped.
// MINW: waves per SIMD the registers must allow (1: no bound; 6: 80 VGPRs,
// no spills; 8: 64 VGPRs with a few spilled)
__device__ inline void scene_tile(const SceneFineArgs& a, TileLDS& L, uint32_t tile) {
  const int lane = threadIdx.x;
  const int TX = (int)(tile % (uint32_t)a.tiles_x), TY = (int)(tile / (uint32_t)a.tiles_x);
  const int col = lane & 15;
  rzp::C4 c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int x = TX * 16 + col, y = TY * 16 + 4 * k + (lane >> 4);
    c[k] = a.load_dst && x < a.lw && y < a.lh ? a.out[(size_t)y * a.row_stride + x] : a.bg;
  }
  const uint32_t nrow = a.row_count[TY];
  const uint32_t* row = a.rows + (size_t)TY * a.n_paths;
  for (uint32_t start = 0; start < nrow; start += 64) {
    const uint32_t j = start + lane;
    uint32_t pid = 0, t = 0, rb = 0, re = 0, eb = 0, ee = 0;
    int bdv = 0, txl = 0, tyl = 0;
    bool draw = false;
    if (j < nrow) {
      pid = row[j];
      const int4 r = a.rects[pid];
      if (r.x <= TX && TX < r.z) {
        const PathInfo& P = a.pi[pid];
        txl = TX - P.tx0;
        tyl = TY - P.ty0;
        t = P.tile_base + (uint32_t)(tyl * P.g.tiles_x + txl);
        rb = a.tile_off[t];
        re = a.tile_off[t + 1];
        eb = a.band_ev_off[P.band_base + tyl];
        ee = a.band_ev_off[P.band_base + tyl + 1];
        bdv = a.bd[t];
        const bool right = txl * 16 - P.g.phx + 15 >= P.redge_col;
        draw = rb != re || eb != ee || bdv != 0 || right;
      }
    }
    uint64_t m = __ballot(draw);
    while (m) {
      const int src = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      const uint32_t q = (uint32_t)__builtin_amdgcn_readlane((int)pid, src);
      const uint32_t qrb = (uint32_t)__builtin_amdgcn_readlane((int)rb, src);
      const uint32_t qre = (uint32_t)__builtin_amdgcn_readlane((int)re, src);
      const uint32_t qeb = (uint32_t)__builtin_amdgcn_readlane((int)eb, src);
      const uint32_t qee = (uint32_t)__builtin_amdgcn_readlane((int)ee, src);
      const int qbd = __builtin_amdgcn_readlane(bdv, src);
      const int qtxl = __builtin_amdgcn_readlane(txl, src);
      const int qtyl = __builtin_amdgcn_readlane(tyl, src);
      const PathInfo& P = a.pi[q];
      const bool local = qrb != qre || qeb != qee;
      const int64_t bd = (int64_t)qbd << rzg::kFixShift;
      if (local) tile_coverage(L, a.recs, qrb, qre, a.events, qeb, qee, qtxl, lane);
      const int xs = qtxl * 16 - P.g.phx + col;  // sub-view pixel column
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int r = 4 * k + (lane >> 4);
        int64_t w = bd;
        if (local) w += tile_local(L, k, lane);
        const int ys = qtyl * 16 - P.g.phy + r;
        if (xs >= 0 && xs < P.g.W && ys >= 0 && ys < P.g.H) {
          if (xs >= P.redge_col) {
            const int64_t rr = a.redge[P.redge_base + ys];
            w += xs == P.redge_col ? (int64_t)rint((double)rr * P.redge_w) : rr;
          }
          const float alpha = P.rule == 0 ? alpha_fix<0>(w) : alpha_fix<1>(w);
          if (alpha >= 1e-6f) {
            const rzp::C4 pc = rzp::paint_at(P.paint, xs, ys);
            c[k] = rzp::blend_over(c[k], rzp::mul(pc, alpha));
          }
        }
      }
      if (local) __syncthreads();  // LDS is reused by the next path
    }
  }
  const int x = TX * 16 + col;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int y = TY * 16 + 4 * k + (lane >> 4);
    if (x < a.lw && y < a.lh) a.out[(size_t)y * a.row_stride + x] = c[k];
  }
}

// One wave per workgroup. Without a queue (QUEUE false), workgroup i takes
// canvas tile i. With one, the grid is persistent (a few waves per SIMD) and
// each wave takes tiles from the front of a.order, costliest first, until
// none is left: the tiles with the most paths start first instead of last.
// The queue head comes as its own restrict parameter: the compiler then
// knows its atomics touch nothing the tiles read, and the tiles' uniform
// loads stay scalar.
// MINW: waves per SIMD the registers must allow (1: no bound; 6: 80 VGPRs;
// 8: 64 VGPRs with some spilled).
template <int MINW, bool QUEUE>
__global__ __launch_bounds__(64, MINW) void k_scene_fine(SceneFineArgs a,
                                                         uint32_t* __restrict__ queue) {
  __shared__ TileLDS L;
  if (a.st->err || a.st->overflow) {
  } else if (!QUEUE) {
    scene_tile(a, L, blockIdx.x);
  } else {
    for (;;) {
      uint32_t t = 0;
      if (threadIdx.x == 0) t = atomicAdd(queue, 1u);
      t = (uint32_t)__shfl((int)t, 0, 64);
      if (t >= a.n_tiles) break;
      __syncthreads();  // the previous tile's LDS reads are done
      scene_tile(a, L, a.order[t]);
    }
  }
  // after the work: a global store ahead of the loads would keep the
  // compiler from reading uniform path data with scalar loads
  if (blockIdx.x == 0 && threadIdx.x == 0 && a.acc) merge_status(a.st, a.acc);
}

// The tile order of the queued form, in one workgroup of kOrderBlock
// threads: each tile's count of paths whose tile rectangle covers it (a 2D
// difference array over the rectangles, then prefix sums along rows and
// columns), tiles sorted by that count, highest first (a counting sort over
// min(count, 1023)), and the queue's head reset.
constexpr int kOrderBlock = 1024;
constexpr int kOrderBuckets = 1024;
__global__ __launch_bounds__(kOrderBlock) void k_scene_order(const int4* rects, uint32_t n_paths,
                                                             int tiles_x, int tiles_y, int* cnt,
                                                             uint32_t* order, uint32_t* queue,
                                                             const Status* st) {
  __shared__ uint32_t hist[kOrderBuckets];
  const int tid = threadIdx.x;
  const uint32_t nt = (uint32_t)tiles_x * (uint32_t)tiles_y;
  if (tid == 0) *queue = 0;
  if (st->err || st->overflow) return;
  // (tiles_x + 1) x (tiles_y + 1) differences
  const uint32_t nd = (uint32_t)(tiles_x + 1) * (uint32_t)(tiles_y + 1);
  for (uint32_t i = tid; i < nd; i += kOrderBlock) cnt[i] = 0;
  for (int i = tid; i < kOrderBuckets; i += kOrderBlock) hist[i] = 0;
  __syncthreads();
  const int W1 = tiles_x + 1;
  for (uint32_t p = tid; p < n_paths; p += kOrderBlock) {
    const int4 r = rects[p];  // [x0, x1) x [y0, y1) in canvas tiles
    const int x0 = min(max(r.x, 0), tiles_x), x1 = min(max(r.z, 0), tiles_x);
    const int y0 = min(max(r.y, 0), tiles_y), y1 = min(max(r.w, 0), tiles_y);
    if (x0 >= x1 || y0 >= y1) continue;
    atomicAdd(&cnt[y0 * W1 + x0], 1);
    atomicAdd(&cnt[y0 * W1 + x1], -1);
    atomicAdd(&cnt[y1 * W1 + x0], -1);
    atomicAdd(&cnt[y1 * W1 + x1], 1);
  }
  __syncthreads();
  for (int y = tid; y < tiles_y; y += kOrderBlock) {
    int run = 0;
    for (int x = 0; x < tiles_x; ++x) cnt[y * W1 + x] = run += cnt[y * W1 + x];
  }
  __syncthreads();
  for (int x = tid; x < tiles_x; x += kOrderBlock) {
    int run = 0;
    for (int y = 0; y < tiles_y; ++y) cnt[y * W1 + x] = run += cnt[y * W1 + x];
  }
  __syncthreads();
  auto bucket = [&](uint32_t t) {
    const int c = cnt[(int)(t / (uint32_t)tiles_x) * W1 + (int)(t % (uint32_t)tiles_x)];
    return (uint32_t)(kOrderBuckets - 1 - min(max(c, 0), kOrderBuckets - 1));  // high counts first
  };
  for (uint32_t t = tid; t < nt; t += kOrderBlock) atomicAdd(&hist[bucket(t)], 1u);
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (int b = 0; b < kOrderBuckets; ++b) {
      const uint32_t v = hist[b];
      hist[b] = run;
      run += v;
    }
  }
  __syncthreads();
  for (uint32_t t = tid; t < nt; t += kOrderBlock) order[atomicAdd(&hist[bucket(t)], 1u)] = t;
}

}  // namespace rzk

namespace rzk {

// ---------------------------------------------------------------------------
// Glyph atlas: Rasterizer::mask of many independent paths, each into its own
// sub-view of one atlas image (src/image.rs:300-311), in one pipeline.
// ---------------------------------------------------------------------------

// One glyph's placement, with its slices of the per-tile / per-band / per-row
// arrays (prefix sums of its tile, band and row counts, made on the host).
struct AtlasSlot {
  double m[6];
  int x, y, w, h;
  uint32_t tile_base, band_base, redge_base, pad;
};

__global__ __launch_bounds__(kBlock) void k_atlas_setup(const AtlasSlot* slots, uint32_t n, int mode,
                                                         PathInfo* pi) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const AtlasSlot& s = slots[i];
  PathInfo P;
  for (int k = 0; k < 6; ++k) P.m[k] = s.m[k];
  P.live = s.w > 0 && s.h > 0;
  P.rule = 0;
  P.ox = s.x;
  P.oy = s.y;
  P.tx0 = P.ty0 = 0;
  P.g.mode = mode;
  P.g.W = s.w;
  P.g.H = s.h;
  P.g.wc = mode == rzg::EDGE_MASK ? (double)s.w - 1.0 : (double)s.w;
  P.g.phx = P.g.phy = 0;
  P.g.tiles_x = (s.w + 15) >> 4;
  P.g.tiles_y = (s.h + 15) >> 4;
  P.tile_base = s.tile_base;
  P.band_base = s.band_base;
  P.redge_base = s.redge_base;
  if (mode == rzg::EDGE_NONE) {
    P.redge_col = INT32_MAX;
    P.redge_w = 0.0;
  } else {
    const double X = rzg::redge_x(P.g), fx = floor(X);
    P.redge_col = (int)fx;
    P.redge_w = 1.0 - (X - fx);
  }
  P.paint = rzp::DPaint{};
  pi[i] = P;
}

// What the fine stage needs of one atlas tile, gathered once per run (after
// the scans) so that the per-tile work starts from a single load.
struct AtlasTile {
  double redge_w;  // the glyph's right-edge weight
  int x, y;        // atlas pixel of the tile's top-left corner
  int w, h;        // glyph pixels inside the tile (0: nothing to draw)
  int txl;         // tile column within the glyph
  int rcol;        // right-edge column relative to the tile's first column
  uint32_t redge;  // right-edge slot of the tile's first row
  uint32_t rb, re;  // the tile's records
  uint32_t eb, ee;  // the events of its band
  int bd;           // its band backdrop
};

// After the scans: one thread per atlas tile.
__global__ __launch_bounds__(kBlock) void k_atlas_tiles(const PathInfo* pi,
                                                         const uint32_t* tile_glyph, uint32_t n_tiles,
                                                         const uint32_t* tile_off,
                                                         const uint32_t* band_ev_off,
                                                         const int32_t* bd, AtlasTile* tiles) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= n_tiles) return;
  const PathInfo& P = pi[tile_glyph[t]];
  AtlasTile T{};
  if (P.live) {
    const uint32_t local = t - P.tile_base;
    const int txl = (int)(local % (uint32_t)P.g.tiles_x), tyl = (int)(local / (uint32_t)P.g.tiles_x);
    T.redge_w = P.redge_w;
    T.x = P.ox + txl * 16;
    T.y = P.oy + tyl * 16;
    T.w = min(16, P.g.W - txl * 16);
    T.h = min(16, P.g.H - tyl * 16);
    T.txl = txl;
    T.rcol = P.redge_col == INT32_MAX ? INT32_MAX : P.redge_col - txl * 16;
    T.redge = P.redge_base + (uint32_t)tyl * 16u;
    const uint32_t band = P.band_base + (uint32_t)tyl;
    T.rb = tile_off[t];
    T.re = tile_off[t + 1];
    T.eb = band_ev_off[band];
    T.ee = band_ev_off[band + 1];
    T.bd = bd[t];
  }
  tiles[t] = T;
}

struct AtlasFineArgs {
  const AtlasTile* tiles;
  uint32_t n;
  const float4* recs;
  const int4* events;
  const long long* redge;
  const Status* st;
  Status* acc;        // accumulated status (merged by block 0)
  float* out;
  size_t row_stride;  // floats
};

// One tile of one glyph per wave (from its AtlasTile); coverage as in k_fine,
// written into the glyph's atlas view.
template <int RULE>
__device__ inline void atlas_tile(const AtlasFineArgs& a, TileLDS& L) {
  const int lane = threadIdx.x;
  const uint32_t t = blockIdx.x;
  const AtlasTile T = a.tiles[t];
  if (T.w <= 0 || T.h <= 0) return;
  const uint32_t rb = T.rb, re = T.re, eb = T.eb, ee = T.ee;
  const int64_t bd = (int64_t)T.bd << rzg::kFixShift;
  const bool local_work = rb != re || eb != ee;
  if (local_work) tile_coverage(L, a.recs, rb, re, a.events, eb, ee, T.txl, lane);
  const int col = lane & 15;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int r = 4 * k + (lane >> 4);
    int64_t w = bd;
    if (local_work) w += tile_local(L, k, lane);
    if (col < T.w && r < T.h) {
      if (col >= T.rcol) {
        const int64_t rr = a.redge[T.redge + (uint32_t)r];
        w += col == T.rcol ? (int64_t)rint((double)rr * T.redge_w) : rr;
      }
      a.out[(size_t)(T.y + r) * a.row_stride + (size_t)(T.x + col)] = alpha_fix<RULE>(w);
    }
  }
}

// The status merge comes after the tile's work (see k_scene_fine).
template <int RULE>
__global__ __launch_bounds__(64) void k_atlas_fine(AtlasFineArgs a) {
  __shared__ TileLDS L;
  if (!(a.st->err || a.st->overflow)) atlas_tile<RULE>(a, L);
  if (blockIdx.x == 0 && threadIdx.x == 0) merge_status(a.st, a.acc);
}

}  // namespace rzk
