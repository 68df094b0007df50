// Adjoint of the batched scene render: dL/d(layer RGBA) -> dL/d(control
// points, solid paint colours, gradient stop colours, transform). No
// counterpart in the reference.
//
// Forward per pixel, paths k in order (Scene::render, src/scene.rs:410-430,
// Rasterizer::fill, src/rasterize.rs:102-115):
//   W_k   = box-filtered winding of path k at the pixel (the signed area that
//           signed_difference_line deposits, src/rasterize.rs:365-470)
//   a_k   = alpha_from_winding(W_k) (src/path.rs:32-46); skipped if < 1e-6
//   s_k   = c_k * a_k                      (LinColor::with_alpha)
//   dst_k = s_k + dst_{k-1} * (1 - s_k.a)  (LinColor::blend_over)
// Backward, with T_k = prod_{j>k} (1 - s_j.a) and g = dL/d dst_N:
//   dL/ds_k = T_k * (g - e_alpha * <g, dst_{k-1}>)
//   dL/da_k = <c_k, dL/ds_k>,  dL/dc_k += a_k * dL/ds_k,  dL/dW_k = dL/da_k * a_k'(W_k)
// W_k(pixel) = integral over the pixel of the winding, which a line (p0, p1)
// changes when it moves: for a closed chain the endpoint terms cancel and
//   dL/dx0 = -I0, dL/dy0 = s I0, dL/dx1 = -I1, dL/dy1 = s I1,
//   I0 = int_{y0}^{y1} g_W(x(y), y) (1 - l(y)) dy, I1 = int g_W l dy,
// with l(y) = (y - y0) / (y1 - y0), s = dx/dy and g_W = dL/dW of the pixel the
// line passes through (piecewise constant, so the integrals are exact sums of
// trapezoids). A line's ends are the curve at t = bits / 2^depth of its
// subdivision leaf, so their gradients reach the control points through the
// Bernstein weights at those t; the subdivision is recomputed, not stored.
#pragma once
#include "rz_scene.hpp"

namespace rzk {

struct BwdTileArgs {
  SceneFineArgs f;              // the forward render's state
  const rzp::C4* grad_out;      // dL/d layer
  size_t grad_stride;           // pixels
  float* wbuf;                  // [entries * 256]: winding W, then dL/dW
  rzp::C4* pbuf;                // [entries * 256]: dst below the path
  float* grad_colors;           // [n_paths * 4]
  float* stop_acc;              // [stops * 4] dL/d(stop as the paint reads it), or null
  const rzp::C4* stop_base;     // the scene's stop table (a paint's stops are a range of it)
};

__device__ inline float alpha_of(int rule, float W) {
  if (rule == 0) {
    const float a = fabsf(W);
    return a >= 1.0f ? 1.0f : (a < 1e-6f ? 0.0f : a);
  }
  const float f = (float)rzp::rem_euclid((double)W + 1.0, 2.0) - 1.0f;
  return fabsf(f);
}

// d alpha / d W where alpha is differentiable (0 on the clamped parts)
__device__ inline float dalpha_of(int rule, float W) {
  if (rule == 0) {
    const float a = fabsf(W);
    return (a >= 1e-6f && a < 1.0f) ? (W < 0.0f ? -1.0f : 1.0f) : 0.0f;
  }
  const float f = (float)rzp::rem_euclid((double)W + 1.0, 2.0) - 1.0f;
  return f < 0.0f ? -1.0f : 1.0f;
}

// Path q over canvas tile (TX, TY): its tile slot, record and event ranges
// and backdrop. Returns whether its winding can be nonzero anywhere on the
// tile (records, events, a backdrop, or pixels at or right of its right-edge
// column); where not, the path contributes nothing to the tile.
__device__ inline bool path_tile(const SceneFineArgs& a, uint32_t q, int TX, int TY, uint32_t* t,
                                 uint32_t* rb, uint32_t* re, uint32_t* eb, uint32_t* ee) {
  const PathInfo& P = a.pi[q];
  const int txl = TX - P.tx0, tyl = TY - P.ty0;
  *t = P.tile_base + (uint32_t)(tyl * P.g.tiles_x + txl);
  *rb = a.tile_off[*t];
  *re = a.tile_off[*t + 1];
  *eb = a.band_ev_off[P.band_base + tyl];
  *ee = a.band_ev_off[P.band_base + tyl + 1];
  const bool right = txl * 16 - P.g.phx + 15 >= P.redge_col;
  return *rb != *re || *eb != *ee || a.bd[*t] != 0 || right;
}

// Per canvas tile: the forward composite again, keeping each path's winding
// and the colour below it; then the paths in reverse order with the
// transmittance of the paths above, giving dL/dW per pixel and path and
// dL/dc per path. Paths with zero winding over the tile (alpha 0 on all of
// it) change neither sweep; only their dL/dW slot is cleared.
__global__ __launch_bounds__(64) void k_scene_bwd_tiles(BwdTileArgs b) {
  const SceneFineArgs& a = b.f;
  if (a.st->err || a.st->overflow) return;
  __shared__ TileLDS L;
  const int lane = threadIdx.x;
  const int TX = (int)(blockIdx.x % (uint32_t)a.tiles_x), TY = (int)(blockIdx.x / (uint32_t)a.tiles_x);
  const int col = lane & 15;
  const uint32_t nrow = a.row_count[TY];
  const uint32_t* row = a.rows + (size_t)TY * a.n_paths;
  rzp::C4 c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) c[k] = a.bg;
  // forward sweep
  for (uint32_t start = 0; start < nrow; start += 64) {
    const uint32_t j = start + lane;
    uint32_t pid = 0;
    bool hit = false;
    if (j < nrow) {
      pid = row[j];
      const int4 r = a.rects[pid];
      hit = r.x <= TX && TX < r.z;
    }
    // a path with zero winding over the tile is skipped (the reverse sweep
    // writes its zero dL/dW)
    uint32_t lt = 0, lrb = 0, lre = 0, leb = 0, lee = 0;
    const bool draw = hit && path_tile(a, pid, TX, TY, &lt, &lrb, &lre, &leb, &lee);
    uint64_t m = __ballot(draw);
    while (m) {
      const int src = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      const uint32_t q = (uint32_t)__builtin_amdgcn_readlane((int)pid, src);
      const PathInfo& P = a.pi[q];
      const int txl = TX - P.tx0, tyl = TY - P.ty0;
      const uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)lt, src);
      const uint32_t rb = (uint32_t)__builtin_amdgcn_readlane((int)lrb, src);
      const uint32_t re = (uint32_t)__builtin_amdgcn_readlane((int)lre, src);
      const uint32_t eb = (uint32_t)__builtin_amdgcn_readlane((int)leb, src);
      const uint32_t ee = (uint32_t)__builtin_amdgcn_readlane((int)lee, src);
      const bool local = rb != re || eb != ee;
      const int64_t bd = (int64_t)a.bd[t] << rzg::kFixShift;
      if (local) tile_coverage(L, a.recs, rb, re, a.events, eb, ee, txl, lane);
      const int xs = txl * 16 - P.g.phx + col;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int r = 4 * k + (lane >> 4);
        int64_t w = bd;
        if (local) w += tile_local(L, k, lane);
        const int ys = tyl * 16 - P.g.phy + r;
        float W = 0.0f;
        const size_t e = (size_t)t * 256 + (size_t)(r * 16 + col);
        if (xs >= 0 && xs < P.g.W && ys >= 0 && ys < P.g.H) {
          if (xs >= P.redge_col) {
            const int64_t rr = a.redge[P.redge_base + ys];
            w += xs == P.redge_col ? (int64_t)rint((double)rr * P.redge_w) : rr;
          }
          W = (float)w * (1.0f / kFix);
          const float alpha = alpha_of(P.rule, W);
          if (alpha >= 1e-6f) {
            // the colour below is needed only where the path is drawn (the
            // reverse sweep tests the same alpha)
            b.pbuf[e] = c[k];
            c[k] = rzp::blend_over(c[k], rzp::mul(rzp::paint_at(P.paint, xs, ys), alpha));
          }
        }
        b.wbuf[e] = W;
      }
      if (local) __syncthreads();
    }
  }
  // reverse sweep
  rzp::C4 g[4];
  float T[4];
  const int x = TX * 16 + col;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int y = TY * 16 + 4 * k + (lane >> 4);
    g[k] = (x < a.lw && y < a.lh) ? b.grad_out[(size_t)y * b.grad_stride + x] : rzp::splat(0.0f);
    T[k] = 1.0f;
  }
  const uint32_t nchunks = (nrow + 63) / 64;
  for (uint32_t ci = nchunks; ci-- > 0;) {
    const uint32_t j = ci * 64 + lane;
    uint32_t pid = 0;
    bool hit = false;
    if (j < nrow) {
      pid = row[j];
      const int4 r = a.rects[pid];
      hit = r.x <= TX && TX < r.z;
    }
    uint32_t lt = 0, lrb, lre, leb, lee;
    const bool draw = hit && path_tile(a, pid, TX, TY, &lt, &lrb, &lre, &leb, &lee);
    // paths skipped by the forward sweep: dL/dW = 0 on their tile slot
    uint64_t z = __ballot(hit && !draw);
    while (z) {
      const int src = __ffsll((unsigned long long)z) - 1;
      z &= z - 1;
      const uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)lt, src);
      reinterpret_cast<float4*>(b.wbuf + (size_t)t * 256)[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    uint64_t m = __ballot(draw);
    // the next path's windings and colours below are loaded while the
    // current one is processed (colours below also where they are unused)
    float nW[4];
    rzp::C4 nB[4];
    auto prefetch = [&](uint64_t mm) {
      if (!mm) return;
      const int s2 = 63 - __clzll((long long)mm);
      const size_t base = (size_t)(uint32_t)__builtin_amdgcn_readlane((int)lt, s2) * 256 + lane;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        nW[k] = b.wbuf[base + 64 * k];
        nB[k] = b.pbuf[base + 64 * k];
      }
    };
    prefetch(m);
    while (m) {
      const int src = 63 - __clzll((long long)m);
      m &= ~(1ull << src);
      float cW[4];
      rzp::C4 cB[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        cW[k] = nW[k];
        cB[k] = nB[k];
      }
      prefetch(m);
      const uint32_t q = (uint32_t)__builtin_amdgcn_readlane((int)pid, src);
      const PathInfo& P = a.pi[q];
      const int txl = TX - P.tx0, tyl = TY - P.ty0;
      const uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)lt, src);
      const int xs = txl * 16 - P.g.phx + col;
      rzp::C4 dc = rzp::splat(0.0f);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int r = 4 * k + (lane >> 4);
        const int ys = tyl * 16 - P.g.phy + r;
        const size_t e = (size_t)t * 256 + (size_t)(r * 16 + col);
        float gW = 0.0f;
        if (xs >= 0 && xs < P.g.W && ys >= 0 && ys < P.g.H) {
          const float W = cW[k];
          const float alpha = alpha_of(P.rule, W);
          if (alpha >= 1e-6f) {
            const rzp::C4 cp = rzp::paint_at(P.paint, xs, ys);
            const rzp::C4 below = cB[k];
            const float gdot = g[k].r * below.r + g[k].g * below.g + g[k].b * below.b + g[k].a * below.a;
            const rzp::C4 ds = rzp::c4(T[k] * g[k].r, T[k] * g[k].g, T[k] * g[k].b, T[k] * (g[k].a - gdot));
            const float da = cp.r * ds.r + cp.g * ds.g + cp.b * ds.b + cp.a * ds.a;
            dc = rzp::add(dc, rzp::mul(ds, alpha));
            gW = da * dalpha_of(P.rule, W);
            int i0, i1;
            float w1;
            rzp::C4 mixed;
            if (b.stop_acc && rzp::gradient_stops_at(P.paint, xs, ys, &i0, &i1, &w1, &mixed)) {
              // through into_linear and the stops' lerp (GradStops::at)
              rzp::C4 gm = rzp::mul(ds, alpha);
              if (!P.paint.linear_colors) gm = rzp::into_linear_vjp(mixed, gm);
              float* o = b.stop_acc + 4 * (size_t)(P.paint.stop_col - b.stop_base);
              const float w0 = 1.0f - w1;
              atomicAdd(o + 4 * i0 + 0, gm.r * w0);
              atomicAdd(o + 4 * i0 + 1, gm.g * w0);
              atomicAdd(o + 4 * i0 + 2, gm.b * w0);
              atomicAdd(o + 4 * i0 + 3, gm.a * w0);
              if (w1 != 0.0f) {
                atomicAdd(o + 4 * i1 + 0, gm.r * w1);
                atomicAdd(o + 4 * i1 + 1, gm.g * w1);
                atomicAdd(o + 4 * i1 + 2, gm.b * w1);
                atomicAdd(o + 4 * i1 + 3, gm.a * w1);
              }
            }
            T[k] *= 1.0f - cp.a * alpha;
          }
        }
        b.wbuf[e] = gW;
      }
      // dL/dc: reduce over the wave, one atomic per channel
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        dc.r += __shfl_xor(dc.r, o, 64);
        dc.g += __shfl_xor(dc.g, o, 64);
        dc.b += __shfl_xor(dc.b, o, 64);
        dc.a += __shfl_xor(dc.a, o, 64);
      }
      // colour gradients are those of solid paints (a gradient paint's colour
      // varies over the path)
      if (lane == 0 && P.paint.kind == rzp::PAINT_SOLID) {
        atomicAdd(&b.grad_colors[4 * q + 0], dc.r);
        atomicAdd(&b.grad_colors[4 * q + 1], dc.g);
        atomicAdd(&b.grad_colors[4 * q + 2], dc.b);
        atomicAdd(&b.grad_colors[4 * q + 3], dc.a);
      }
    }
  }
}

// dL/dW of the path's pixel (x, y) of its sub-view, 0 outside the image.
__device__ inline float gW_at(const PathInfo& P, const float* wbuf, int x, int y) {
  if (x < 0 || y < 0 || x >= P.g.W || y >= P.g.H) return 0.0f;
  const int gx = x + P.g.phx, gy = y + P.g.phy;
  const uint32_t t = P.tile_base + (uint32_t)((gy >> 4) * P.g.tiles_x + (gx >> 4));
  return wbuf[(size_t)t * 256 + (size_t)((gy & 15) * 16 + (gx & 15))];
}

// I0, I1 of one line (sub-view frame): for each pixel row it crosses and each
// pixel column of that row, g_W times the exact integral of (1 - l) and l
// over the y-range of the line inside the pixel.
__device__ inline void line_integrals(const PathInfo& P, const float* wbuf, Pt a, Pt c, double* I0,
                                      double* I1) {
  *I0 = *I1 = 0.0;
  const double dy = c.y - a.y;
  if (dy == 0.0) return;
  const double dxdy = (c.x - a.x) / dy;
  const double ylo = fmax(fmin(a.y, c.y), 0.0), yhi = fmin(fmax(a.y, c.y), (double)P.g.H);
  if (!(ylo < yhi)) return;
  const int r0 = (int)floor(ylo), r1 = (int)ceil(yhi);
  const double sgn = dy > 0.0 ? 1.0 : -1.0;
  for (int r = r0; r < r1; ++r) {
    const double ya = fmax(ylo, (double)r), yb = fmin(yhi, (double)r + 1.0);
    if (!(ya < yb)) continue;
    const double xa = a.x + (ya - a.y) * dxdy, xb = a.x + (yb - a.y) * dxdy;
    const int cmin = max((int)floor(fmin(xa, xb)), 0);
    const int cmax = min((int)floor(fmax(xa, xb)), P.g.W - 1);
    for (int cx = cmin; cx <= cmax; ++cx) {
      double lo = ya, hi = yb;
      if (dxdy != 0.0) {
        const double y0 = a.y + ((double)cx - a.x) / dxdy, y1 = a.y + ((double)cx + 1.0 - a.x) / dxdy;
        lo = fmax(lo, fmin(y0, y1));
        hi = fmin(hi, fmax(y0, y1));
      }
      if (!(lo < hi)) continue;
      const float gw = gW_at(P, wbuf, cx, r);
      if (gw == 0.0f) continue;
      const double l0 = (lo - a.y) / dy, l1 = (hi - a.y) / dy;
      const double len = sgn * (hi - lo);  // signed, in the line's direction
      const double lm = 0.5 * (l0 + l1);
      *I0 += (double)gw * len * (1.0 - lm);
      *I1 += (double)gw * len * lm;
    }
  }
}

struct BwdLineArgs {
  const PathInfo* pi;
  const uint32_t* slot_path;
  const uint32_t* subpaths;  // combined path: subpath -> first segment, [n_sub + 1]
  const float* wbuf;
  float* grad_points;        // [2 * points] of the combined path
  double* grad_tr_part;      // [waves * 6] per-wave partial dL/d transform
};

// Bernstein weights of a NP-point Bezier at t.
template <int NP>
__device__ inline void bernstein(double t, double* w) {
  const double u = 1.0 - t;
  if (NP == 2) {
    w[0] = u;
    w[1] = t;
  } else if (NP == 3) {
    w[0] = u * u;
    w[1] = 2.0 * t * u;
    w[2] = t * t;
  } else {
    w[0] = u * u * u;
    w[1] = 3.0 * t * u * u;
    w[2] = 3.0 * t * t * u;
    w[3] = t * t * t;
  }
}

// One flatten item of the backward pass: the same subtree as k_flat_lines, each
// leaf line's endpoint gradients pushed to the slot's control points.
template <int NP>
__device__ void bwd_line_item(const DPath& p, uint32_t slot, uint32_t j, int D0, bool active,
                              const Xform& x, double fl, const PathInfo* pi,
                              const uint32_t* slot_path, const BwdLineArgs& b, double* gtr) {
  rzg::Node<NP> n;
  bool walk, leaf;
  int err;
  uint32_t pid;
  flat_subtree<NP, true>(p, slot, j, D0, active, x, fl, pi, slot_path, &n, &walk, &leaf, &err, &pid);
  if (!walk && !leaf) return;
  const PathInfo& P = pi[pid];
  double gp[NP][2];
#pragma unroll
  for (int i = 0; i < NP; ++i) gp[i][0] = gp[i][1] = 0.0;
  auto on_leaf = [&](Pt a, Pt c, int depth, uint64_t bits) {
    double I0, I1;
    line_integrals(P, b.wbuf, a, c, &I0, &I1);
    if (I0 == 0.0 && I1 == 0.0) return;
    const double s = (c.x - a.x) / (c.y - a.y);
    const double ga[2] = {-I0, s * I0}, gc[2] = {-I1, s * I1};
    const double scale = ldexp(1.0, -depth);
    double wa[NP], wc[NP];
    bernstein<NP>((double)bits * scale, wa);
    bernstein<NP>((double)(bits + 1) * scale, wc);
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      gp[i][0] += wa[i] * ga[0] + wc[i] * gc[0];
      gp[i][1] += wa[i] * ga[1] + wc[i] * gc[1];
    }
  };
  if (leaf) {
    // a leaf above the subtree depth: its bits are the top bits of j
    int k = 0;
    uint64_t bits = 0;
    // find the leaf depth again (the walk stopped where the node was flat)
    rzg::Node<NP> r = load_node<NP>(p, slot, P.m);
    while (k < D0 && !(rzg::flatness(r) < fl)) {
      rzg::Node<NP> ch;
      const int bit = (int)((j >> (D0 - 1 - k)) & 1u);
      rzg::child(r, bit, &ch);
      r = ch;
      bits = (bits << 1) | (uint64_t)bit;
      ++k;
    }
    on_leaf(n.p[0], n.p[NP - 1], k, bits);
  } else {
    rzg::traverse_db(n, fl, [&](Pt a, Pt c, int depth, uint64_t bits) {
      on_leaf(a, c, D0 + depth, ((uint64_t)j << depth) | bits);
    });
  }
  // control point indices of the slot (a closing line joins the subpath's
  // last point to its first)
  uint32_t idx[NP];
  const uint32_t item = p.slot_item[slot];
  if (item & 0x80000000u) {
    const uint32_t sub = item & 0x7fffffffu;
    const uint32_t s0 = b.subpaths[sub], s1 = b.subpaths[sub + 1];
    idx[0] = p.pt_off[s1 - 1] + (uint32_t)p.kinds[s1 - 1] + 1u;
    idx[NP - 1] = p.pt_off[s0];
  } else {
#pragma unroll
    for (int i = 0; i < NP; ++i) idx[i] = p.pt_off[item] + (uint32_t)i;
  }
  // transformed point p' = m p: dL/dp = M^T dL/dp', dL/dm += dL/dp' (x, y, 1)
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const double gx = gp[i][0], gy = gp[i][1];
    if (gx == 0.0 && gy == 0.0) continue;
    const double px = p.pts[2 * (size_t)idx[i]], py = p.pts[2 * (size_t)idx[i] + 1];
    atomicAdd(&b.grad_points[2 * (size_t)idx[i]], (float)(P.m[0] * gx + P.m[3] * gy));
    atomicAdd(&b.grad_points[2 * (size_t)idx[i] + 1], (float)(P.m[1] * gx + P.m[4] * gy));
    gtr[0] += gx * px;
    gtr[1] += gx * py;
    gtr[2] += gx;
    gtr[3] += gy * px;
    gtr[4] += gy * py;
    gtr[5] += gy;
  }
}

template <int NP>
__device__ inline void bwd_line_dispatch(const DPath& p, uint32_t slot, uint32_t j, int D0,
                                         bool active, const Xform& x, double fl,
                                         const PathInfo* pi, const uint32_t* slot_path,
                                         const BwdLineArgs& b, double* gtr) {
  bwd_line_item<NP>(p, slot, j, D0, active, x, fl, pi, slot_path, b, gtr);
}

__global__ __launch_bounds__(kBlock) void k_scene_bwd_lines(DPath p, Xform x, double fl, uint32_t bl,
                                                             uint32_t bq, int dq, int dc,
                                                             const PathInfo* pi,
                                                             const uint32_t* slot_path,
                                                             BwdLineArgs b, const Status* st) {
  if (st->err || st->overflow) return;
  double gtr[6] = {0, 0, 0, 0, 0, 0};
  const uint32_t blk = blockIdx.x, t = threadIdx.x;
  if (blk < bl) {
    const uint32_t slot = p.order[blk * kBlock + t];
    bwd_line_dispatch<2>(p, slot, 0, 0, slot != ~0u, x, fl, pi, slot_path, b, gtr);
  } else if (blk < bl + bq) {
    const uint32_t v = (blk - bl) * kBlock + t;
    const uint32_t slot = p.order[p.blocks_line * kBlock + (v >> dq)];
    bwd_line_dispatch<3>(p, slot, v & ((1u << dq) - 1u), dq, slot != ~0u, x, fl, pi, slot_path, b,
                         gtr);
  } else {
    const uint32_t v = (blk - bl - bq) * kBlock + t;
    const uint32_t slot = p.order[(p.blocks_line + p.blocks_quad) * kBlock + (v >> dc)];
    bwd_line_dispatch<4>(p, slot, v & ((1u << dc) - 1u), dc, slot != ~0u, x, fl, pi, slot_path, b,
                         gtr);
  }
  // per-wave partial of dL/d transform
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double v = gtr[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    gtr[k] = v;
  }
  if ((threadIdx.x & 63) == 0) {
    const size_t wave = ((size_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
#pragma unroll
    for (int k = 0; k < 6; ++k) b.grad_tr_part[wave * 6 + k] = gtr[k];
  }
}

// Sum the per-wave transform partials (one block). A run that failed or
// outgrew its buffers wrote no partials, and adds nothing.
__global__ __launch_bounds__(kBlock) void k_sum_tr(const double* part, size_t nwaves, double* out,
                                                    const Status* st) {
  if (st->err || st->overflow) return;
  __shared__ double sh[kBlock];
  for (int k = 0; k < 6; ++k) {
    double v = 0.0;
    for (size_t i = threadIdx.x; i < nwaves; i += kBlock) v += part[i * 6 + k];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] += sh[0];
    __syncthreads();
  }
}

}  // namespace rzk
