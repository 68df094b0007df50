// The stroke kernels: Path::stroke (src/path.rs:370-415) over work items
// (one per segment and direction, rz_stroke.hpp), in four passes:
//
//   k_stroke_pieces  each item's offset curves: count, points, first and
//                    last curve; an in-block max-scan of the indices of the
//                    items that have curves (k_stroke_blocks carries it
//                    across blocks)
//   k_stroke_runs    each run's first item with curves
//   k_stroke_count   each item's output: the join in front of its curves,
//                    the curves, and at a run's end its closing or cap
//   k_stroke_emit    the same segments written at the prefix sums of those
//                    counts (rz_path arrays), and each run's subpath end
#pragma once
#include "rz_kernels.hpp"
#include "rz_stroke.hpp"

namespace rzk {

constexpr int kStrokeBlock = 256;

struct StrokeBufs {
  rzs::Seg* first;      // [n_items] first offset curve of an item
  rzs::Seg* last;       // [n_items] last offset curve
  uint32_t* pieces;     // [n_items] offset curves of an item
  uint32_t* piece_pts;  // [n_items] their points
  int32_t* incl;        // [n_items] in-block inclusive max of item-with-curves indices
  int32_t* blk;         // [n_blocks] block maxima, then the exclusive prefix maxima
  int32_t* run_first;   // [n_runs] first item with curves of a run, -1 if none
  uint32_t* run_has;    // [n_runs] whether a run has curves (0/1)
  uint32_t* cnt_seg;    // [n_items] output segments of an item
  uint32_t* cnt_pts;    // [n_items] output points of an item
};

// The nearest item <= g with curves (-1 if none).
__device__ inline int32_t stroke_prev_incl(const StrokeBufs& b, uint32_t g) {
  return max(b.blk[g / kStrokeBlock], b.incl[g]);
}
// The nearest item < g with curves (-1 if none).
__device__ inline int32_t stroke_prev(const StrokeBufs& b, uint32_t g) {
  return g == 0 ? -1 : stroke_prev_incl(b, g - 1);
}

__global__ __launch_bounds__(kStrokeBlock) void k_stroke_pieces(rzs::Input in, StrokeBufs b) {
  __shared__ int32_t wmax[kStrokeBlock / 64];
  const uint32_t g = blockIdx.x * kStrokeBlock + threadIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int32_t v = -1;
  if (g < in.n_items) {
    const rzs::Item it = rzs::item_of(in, g);
    rzs::CountSink c;
    rzs::seg_offset(rzs::item_seg(in, it), in.st.width / 2.0, c);
    b.pieces[g] = c.n;
    b.piece_pts[g] = c.pts;
    if (c.n) {
      b.first[g] = c.first;
      b.last[g] = c.last;
      v = (int32_t)g;
    }
  }
  int32_t x = dpp_max_scan(v);
  if (lane == 63) wmax[wid] = x;
  __syncthreads();
  for (int w = 0; w < wid; ++w) x = max(x, wmax[w]);
  if (g < in.n_items) b.incl[g] = x;
  if (threadIdx.x == kStrokeBlock - 1) b.blk[blockIdx.x] = x;
}

// Exclusive prefix maxima of the block maxima, in place (one block, chunks
// of kBlock with a carry).
__global__ __launch_bounds__(kBlock) void k_stroke_blocks(int32_t* blk, uint32_t n) {
  __shared__ int32_t s[kBlock];
  int32_t carry = -1;
  for (uint32_t base = 0; base < n; base += kBlock) {
    const uint32_t i = base + threadIdx.x;
    const int32_t v = i < n ? blk[i] : -1;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {
      const int32_t y = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : -1;
      __syncthreads();
      s[threadIdx.x] = max(s[threadIdx.x], y);
      __syncthreads();
    }
    const int32_t excl = max(carry, threadIdx.x ? s[threadIdx.x - 1] : -1);
    if (i < n) blk[i] = excl;
    carry = max(carry, s[kBlock - 1]);
    __syncthreads();
  }
}

__global__ __launch_bounds__(kStrokeBlock) void k_stroke_runs(rzs::Input in, StrokeBufs b) {
  const uint32_t g = blockIdx.x * kStrokeBlock + threadIdx.x;
  if (g >= in.n_items || b.pieces[g] == 0) return;
  const rzs::Item it = rzs::item_of(in, g);
  if (stroke_prev(b, g) < (int32_t)it.run_b) {
    b.run_first[it.run] = (int32_t)g;
    b.run_has[it.run] = 1;
  }
}

// The pieces pass's results as rzs::item_output asks for them.
struct StrokeQ {
  const StrokeBufs& b;
  __device__ uint32_t pieces(uint32_t g) const { return b.pieces[g]; }
  __device__ int32_t prev(uint32_t g) const { return stroke_prev(b, g); }
  __device__ int32_t prev_incl(uint32_t g) const { return stroke_prev_incl(b, g); }
  __device__ bool run_has(uint32_t r) const { return b.run_has[r] != 0; }
  __device__ int32_t run_first(uint32_t r) const { return b.run_first[r]; }
  __device__ const rzs::Seg& first(int32_t i) const { return b.first[i]; }
  __device__ const rzs::Seg& last(int32_t i) const { return b.last[i]; }
};

__global__ __launch_bounds__(kStrokeBlock) void k_stroke_count(rzs::Input in, StrokeBufs b) {
  const uint32_t g = blockIdx.x * kStrokeBlock + threadIdx.x;
  if (g >= in.n_items) return;
  const rzs::Item it = rzs::item_of(in, g);
  rzs::CountSink c;
  rzs::item_output(in, StrokeQ{b}, g, it, false, c);
  b.cnt_seg[g] = c.n + b.pieces[g];
  b.cnt_pts[g] = c.pts + b.piece_pts[g];
}

struct StrokeOut {
  const uint32_t* seg_off;  // [n_items] exclusive prefix sums of cnt_seg
  const uint32_t* pt_off;   // [n_items] of cnt_pts
  const uint32_t* run_idx;  // [n_runs] exclusive prefix sums of run_has
  uint8_t* kinds;
  uint32_t* seg_pts;        // point offset of each output segment
  double* pts;
  uint32_t* subs;           // [n_out_subpaths + 1]
  uint8_t* closed;
};

__global__ __launch_bounds__(kStrokeBlock) void k_stroke_emit(rzs::Input in, StrokeBufs b,
                                                             StrokeOut o) {
  const uint32_t g = blockIdx.x * kStrokeBlock + threadIdx.x;
  if (g >= in.n_items) return;
  const rzs::Item it = rzs::item_of(in, g);
  rzs::WriteSink w;
  w.kinds = o.kinds;
  w.pt_off = o.seg_pts;
  w.pts = o.pts;
  w.seg = o.seg_off[g];
  w.pt = o.pt_off[g];
  rzs::item_output(in, StrokeQ{b}, g, it, true, w);
  if (g == it.run_e - 1 && b.run_has[it.run]) {
    const uint32_t k = o.run_idx[it.run];
    o.subs[k + 1] = w.seg;
    o.closed[k] = 1;
  }
}

}  // namespace rzk
