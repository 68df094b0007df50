// Geometry shared by the HIP kernels (device) and the host side: f64 transform
// and segment arithmetic for flattening, and the binning geometry that turns a
// flattened line into per-tile pieces.
//
// Flattening must reproduce the reference's line list bit for bit, so every
// expression below keeps the reference's operation order and the library is
// compiled with -ffp-contract=off (Rust never fuses multiply-add).
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define RZ_HD __host__ __device__ inline
#define RZ_UNROLL _Pragma("unroll")
#else
#define RZ_HD inline
#define RZ_UNROLL
#endif

namespace rzg {

// ---------------------------------------------------------------------------
// Constants (src/geometry.rs:14-16)
// ---------------------------------------------------------------------------
static constexpr double kEpsilon = 2.220446049250313e-16;  // f64::EPSILON
static constexpr int kTile = 16;                            // tile edge in pixels
static constexpr int kFixShift = 20;                        // LDS fixed point: 2^-20
static constexpr int kMaxDepth = 62;
#ifndef RZ_STACK_Q
#define RZ_STACK_Q 4  // register stack entries of the quad traversal
#endif
#ifndef RZ_STACK_C
#define RZ_STACK_C 4  // register stack entries of the cubic traversal
#endif                        // subdivision depth limit

enum SegKind { KIND_LINE = 0, KIND_QUAD = 1, KIND_CUBIC = 2 };
enum EdgeMode { EDGE_MASK = 0, EDGE_ITER = 1, EDGE_NONE = 2 };
enum Status { ST_OK = 0, ST_NAN = 1, ST_DEPTH = 2, ST_WRAP = 3 };

struct Pt {
  double x, y;
};
RZ_HD Pt mkpt(double x, double y) {
  Pt p;
  p.x = x;
  p.y = y;
  return p;
}

// Transform::apply (src/geometry.rs:363-367): x*m00 + y*m01 + m02
RZ_HD Pt tr_apply(const double* m, Pt p) {
  return mkpt(p.x * m[0] + p.y * m[1] + m[2], p.x * m[3] + p.y * m[4] + m[5]);
}

// A Bezier node: K+1 control points (K = 1 line, 2 quad, 3 cubic).
template <int NP>
struct Node {
  Pt p[NP];
};

// Segment::flatness (src/curve.rs:233-235, 413-417, 692-697)
RZ_HD double flatness(const Node<2>&) { return 0.0; }
RZ_HD double flatness(const Node<3>& n) {
  double x = 2.0 * n.p[1].x - n.p[0].x - n.p[2].x;
  double y = 2.0 * n.p[1].y - n.p[0].y - n.p[2].y;
  return x * x + y * y;
}
RZ_HD double flatness(const Node<4>& n) {
  double ux = 3.0 * n.p[1].x - 2.0 * n.p[0].x - n.p[3].x;
  double uy = 3.0 * n.p[1].y - 2.0 * n.p[0].y - n.p[3].y;
  double vx = 3.0 * n.p[2].x - n.p[0].x - 2.0 * n.p[3].x;
  double vy = 3.0 * n.p[2].y - n.p[0].y - 2.0 * n.p[3].y;
  return fmax(ux * ux, vx * vx) + fmax(uy * uy, vy * vy);
}

template <int NP>
RZ_HD bool has_nans(const Node<NP>& n) {
  bool r = false;
  for (int i = 0; i < NP; ++i) r = r || isnan(n.p[i].x) || isnan(n.p[i].y);
  return r;
}

// True when every coordinate is below 1e300 in magnitude (false for NaN).
// Every point of a split is a combination of its parent's points with
// non-negative power-of-two weights summing to 1, so the descendants of such
// a node stay below 1e300 as well (no overflow to infinity, hence no NaN):
// the per-node NaN test of the traversal is needed only for other roots.
template <int NP>
RZ_HD bool bounded(const Node<NP>& n) {
  bool r = true;
  for (int i = 0; i < NP; ++i) r = r && fabs(n.p[i].x) < 1e300 && fabs(n.p[i].y) < 1e300;
  return r;
}

// Curve::split_at(0.5), the split PathFlattenIter uses (src/curve.rs:43-45,
// 1186-1201). With t = 0.5 every coefficient the reference computes
// (1-t, t*t, 2*t*(1-t), 3*t*(1-t)^2, ...) is an exact power-of-two multiple,
// so the literals below are bit-identical to the reference's products; the
// additions keep its left-to-right order (src/curve.rs:458-468, 749-769).
RZ_HD void child(const Node<3>& n, int right, Node<3>* c) {
  Pt mid = mkpt(0.25 * n.p[0].x + 0.5 * n.p[1].x + 0.25 * n.p[2].x,
                0.25 * n.p[0].y + 0.5 * n.p[1].y + 0.25 * n.p[2].y);
  if (!right) {
    c->p[0] = n.p[0];
    c->p[1] = mkpt(0.5 * n.p[0].x + 0.5 * n.p[1].x, 0.5 * n.p[0].y + 0.5 * n.p[1].y);
    c->p[2] = mid;
  } else {
    c->p[0] = mid;
    c->p[1] = mkpt(0.5 * n.p[1].x + 0.5 * n.p[2].x, 0.5 * n.p[1].y + 0.5 * n.p[2].y);
    c->p[2] = n.p[2];
  }
}
RZ_HD void child(const Node<4>& n, int right, Node<4>* c) {
  Pt mid = mkpt(0.125 * n.p[0].x + 0.375 * n.p[1].x + 0.375 * n.p[2].x + 0.125 * n.p[3].x,
                0.125 * n.p[0].y + 0.375 * n.p[1].y + 0.375 * n.p[2].y + 0.125 * n.p[3].y);
  if (!right) {
    c->p[0] = n.p[0];
    c->p[1] = mkpt(0.5 * n.p[0].x + 0.5 * n.p[1].x, 0.5 * n.p[0].y + 0.5 * n.p[1].y);
    c->p[2] = mkpt(0.25 * n.p[0].x + 0.5 * n.p[1].x + 0.25 * n.p[2].x,
                   0.25 * n.p[0].y + 0.5 * n.p[1].y + 0.25 * n.p[2].y);
    c->p[3] = mid;
  } else {
    c->p[0] = mid;
    c->p[1] = mkpt(0.25 * n.p[1].x + 0.5 * n.p[2].x + 0.25 * n.p[3].x,
                   0.25 * n.p[1].y + 0.5 * n.p[2].y + 0.25 * n.p[3].y);
    c->p[2] = mkpt(0.5 * n.p[2].x + 0.5 * n.p[3].x, 0.5 * n.p[2].y + 0.5 * n.p[3].y);
    c->p[3] = n.p[3];
  }
}
RZ_HD void child(const Node<2>& n, int right, Node<2>* c) { *c = n; }

// In-order traversal of one segment's subdivision tree (PathFlattenIter,
// src/path.rs:758-796): a node whose flatness is below `fl` (= 16*tol^2) is a
// leaf and emits Line(start, end); otherwise it is split at t = 0.5 and its
// left child is visited first.
//
// The reference keeps an explicit stack. Here the pending right siblings live
// in a register shift stack of kStack(NP) entries; when that underflows the
// next node is recomputed from the root along its path bits (depth splits
// from the root are the same operations the recursive split performs, so the
// recomputed node is bit-identical). This keeps the traversal in registers
// with no scratch memory; recomputation only happens above the deepest
// kStack(NP) levels. A stack entry leaves out the right sibling's first
// point: when it is popped, the leaf just emitted is the rightmost one under
// the left sibling, whose last point is that sibling's last point (a right
// child copies its parent's last point), i.e. the parent's midpoint, computed
// by the same expression as the right sibling's first point.
// leaf(start, end, depth, bits): the leaf is the root's subtree at `depth`
// reached along `bits` (most significant first, 1 = right child), i.e. the
// curve on t in [bits, bits + 1] / 2^depth of the root.
template <int NP>
constexpr int kStack() { return RZ_STACK_Q * (NP == 3) + RZ_STACK_C * (NP == 4) + 4 * (NP == 2); }

template <int NP>
struct Tail {  // a node without its first point
  Pt p[NP - 1];
};

template <int NP, typename Leaf>
RZ_HD int traverse_db(const Node<NP>& root, double fl, Leaf&& leaf) {
  constexpr int S = kStack<NP>();
  Node<NP> cur = root;
  Tail<NP> st[S];
  int nst = 0, depth = 0;
  uint64_t bits = 0;
  const bool check = !bounded(root);
  while (true) {
    if (check && has_nans(cur)) return ST_NAN;
    if (flatness(cur) < fl) {
      leaf(cur.p[0], cur.p[NP - 1], depth, bits);
      while (depth > 0 && (bits & 1)) {
        bits >>= 1;
        --depth;
      }
      if (depth == 0) return ST_OK;
      bits |= 1;
      if (nst > 0) {
        cur.p[0] = cur.p[NP - 1];
RZ_UNROLL
        for (int i = 1; i < NP; ++i) cur.p[i] = st[0].p[i - 1];
RZ_UNROLL
        for (int k = 0; k + 1 < S; ++k) st[k] = st[k + 1];
        --nst;
      } else {
        Node<NP> n = root;
        for (int k = depth - 1; k >= 0; --k) {
          Node<NP> c;
          child(n, (int)((bits >> k) & 1), &c);
          n = c;
        }
        cur = n;
      }
    } else {
      if (depth >= kMaxDepth) return ST_DEPTH;
      Node<NP> l, r;
      child(cur, 0, &l);
      child(cur, 1, &r);
RZ_UNROLL
      for (int k = S - 1; k > 0; --k) st[k] = st[k - 1];
RZ_UNROLL
      for (int i = 1; i < NP; ++i) st[0].p[i - 1] = r.p[i];
      nst = nst < S ? nst + 1 : S;
      cur = l;
      ++depth;
      bits <<= 1;
    }
  }
}

template <int NP, typename Leaf>
RZ_HD int traverse(const Node<NP>& root, double fl, Leaf&& leaf) {
  return traverse_db(root, fl, [&](Pt a, Pt b, int, uint64_t) { leaf(a, b); });
}

// ---------------------------------------------------------------------------
// Binning geometry
// ---------------------------------------------------------------------------
struct BinGrid {
  double wc;      // right clip abscissa: W-1 (mask), W (mask_iter/fill); unused for EDGE_NONE
  int mode;       // EdgeMode
  int W, H;       // image size in pixels
  int tiles_x, tiles_y;
  // Grid phase: tile column k spans x in [16k - phx, 16k + 16 - phx) and band
  // k spans y in [16k - phy, ...). Zero for a whole image; a scene path drawn
  // into a sub-view at (ox, oy) uses (ox mod 16, oy mod 16) so that its tiles
  // are the canvas tiles.
  int phx, phy;
};

// x of the line (a, b) at ordinate Y, y of it at abscissa X. Both use the
// reference's lerp form (1-t)*p0 + t*p1 (src/rasterize.rs:378), which returns
// the endpoint exactly at t = 0 and t = 1, so pieces that meet on a tile or
// band boundary share their endpoint bit for bit.
RZ_HD Pt at_y(Pt a, Pt b, double Y) {
  double t = (Y - a.y) / (b.y - a.y);
  return mkpt((1.0 - t) * a.x + t * b.x, Y);
}
RZ_HD Pt at_x(Pt a, Pt b, double X) {
  double t = (X - a.x) / (b.x - a.x);
  return mkpt(X, (1.0 - t) * a.y + t * b.y);
}

// The reference's per-line preprocessing inside signed_difference_line: right
// clip at x = wc (src/rasterize.rs:370-387), then the left split at x = 0
// (src/rasterize.rs:389-393, 922-937) whose x<0 part becomes a vertical line on
// x = 0. Produces up to two pieces in `out` (returns the count) and reports
// which endpoints of the effective polyline were created by the right clip.
// A line starting exactly on x = wc and ending to its right never gets here:
// the flatten replaces it by wrap_lines (below).
struct EffLine {
  Pt p[3];   // polyline vertices p[0] -> ... -> p[n]
  int n;     // number of pieces (0, 1 or 2)
  bool clip0, clip1;  // start / end vertex created by the right clip
  bool proj;  // both ends beyond the clip: a vertical line on x = redge_x(g)
};

// Abscissa of the vertical line that replaces a line lying entirely beyond the
// right clip (src/rasterize.rs:372-375).
RZ_HD double redge_x(const BinGrid& g) { return g.wc - 0.001; }

// A line that starts exactly on the right clip abscissa wc and ends to its
// right. signed_difference_line keeps such a line whole: its `p0.x() < width`
// test fails at equality, so the part it keeps is the one beyond the clip
// (src/rasterize.rs:373-384). Every row's deposit pair then lands at columns
// >= wc, and those past the end of the accumulation buffer's row (of `stride`
// entries: W + 1 for mask_iter, W for a dense mask, the row stride for a mask
// into a wider image) continue in the next row: with stride wc + 1, a deposit
// at column wc + 1 is the next row's first pixel.
struct WrapClip {
  double wc;      // the clip abscissa; negative: no such lines (ActiveEdge, atlases)
  double stride;  // entries per buffer row
  int W, H;       // visible columns and rows of the buffer
};
RZ_HD bool wraps(Pt p0, Pt p1, const WrapClip& q) {
  return q.wc >= 0.0 && p0.x == q.wc && p1.x > q.wc;
}

// The flatten stores a line of wrap_lines (below) with this NaN as its x1: a
// payload no arithmetic makes, so the count pass knows the line (vertical on
// its x0) and gives both its ends events, as for the ends the right clip
// makes. These lines are not part of a closed chain.
constexpr uint64_t kWrapTag = 0x7ffc0ffee0000001ull;
RZ_HD double wrap_tag() { return __builtin_bit_cast(double, kWrapTag); }
RZ_HD bool is_wrap_tag(double x) { return __builtin_bit_cast(uint64_t, x) == kWrapTag; }

// The deposits of such a line (the reference's row loop, src/rasterize.rs:
// 398-444, whose single-pixel case is the only one here: x1 is clamped to wc
// and x0 >= wc), each made by a vertical line on its pixel's left edge:
// f(a, b) with a = (c, r) and b = (c, r + v) (reversed when v < 0), which
// deposits v at column c of row r (|v| > 1: several such lines). Deposits
// outside the visible columns (the dead column of mask_iter, c = W, or the
// rest of a wider image's row) or past the last row are dropped. A deposit
// above kWrapSplit in size (a line running more than about 2 kWrapSplit
// pixels right within one row) is refused: ST_WRAP, and nothing past it.
constexpr double kWrapSplit = 256.0;
template <typename F>
RZ_HD int wrap_lines(Pt p0, Pt p1, const WrapClip& q, F&& f) {
  if (fabs(p0.y - p1.y) < kEpsilon) return ST_OK;
  double dir = 1.0;
  if (p0.y > p1.y) {
    dir = -1.0;
    const Pt t = p0;
    p0 = p1;
    p1 = t;
  }
  auto put = [&](double y, double c, double v) {
    const double k = floor(c / q.stride);
    const double r = y + k, cc = c - k * q.stride;
    if (r >= (double)q.H || cc >= (double)q.W || v == 0.0) return true;
    // a row's deposit can exceed 1 in size (xmf > 1 when the line crosses
    // several columns within the row): then n lines of |v| / n each
    const double n = ceil(fabs(v)), h = fabs(v) / n;
    if (!(n <= kWrapSplit)) return false;
    const Pt a = mkpt(cc, r), b = mkpt(cc, r + h);
    for (double i = 0.0; i < n; i += 1.0) {
      if (v > 0.0) f(a, b);
      else f(b, a);
    }
    return true;
  };
  const double dxdy = (p1.x - p0.x) / (p1.y - p0.y);
  double y = floor(fmax(p0.y, 0.0));
  double x = p0.y < 0.0 ? p0.x - p0.y * dxdy : p0.x;
  double x_next = x;
  const double yend = fmin((double)q.H, fmax(ceil(p1.y), 0.0));
  for (; y < yend; y += 1.0) {
    x = x_next;
    const double dy = fmin(y + 1.0, p1.y) - fmax(y, p0.y);
    const double d = dir * dy;
    x_next = x + dxdy * dy;
    const double x0 = x < x_next ? x : x_next;
    const double x0f = fmax(floor(x0), 0.0);
    const double xmf = 0.5 * (x + x_next) - x0f;
    if (!put(y, x0f, d * (1.0 - xmf)) || !put(y, x0f + 1.0, d * xmf)) return ST_WRAP;
  }
  return ST_OK;
}

RZ_HD EffLine effective_line(Pt p0, Pt p1, const BinGrid& g) {
  EffLine e;
  e.clip0 = e.clip1 = false;
  e.proj = false;
  e.n = 1;
  if (g.mode != EDGE_NONE) {
    const double width = g.wc;
    if (p0.x > width || p1.x > width) {
      if (p0.x > width && p1.x > width) {
        p0 = mkpt(redge_x(g), p0.y);
        p1 = mkpt(redge_x(g), p1.y);
        e.clip0 = e.clip1 = true;
        e.proj = true;
      } else {
        double t = (p0.x - width) / (p0.x - p1.x);
        Pt mid = mkpt(width, (1.0 - t) * p0.y + t * p1.y);
        if (p0.x <= width) {
          p1 = mid;
          e.clip1 = true;
        } else {
          p0 = mid;
          e.clip0 = true;
        }
      }
    }
  }
  if (p0.x >= 0.0 && p1.x >= 0.0) {
    e.p[0] = p0;
    e.p[1] = p1;
  } else if (p0.x <= 0.0 && p1.x <= 0.0) {
    e.p[0] = mkpt(0.0, p0.y);
    e.p[1] = mkpt(0.0, p1.y);
  } else {
    double t = p0.x / (p0.x - p1.x);
    Pt mid = mkpt((1.0 - t) * p0.x + t * p1.x, (1.0 - t) * p0.y + t * p1.y);
    if (p0.x < 0.0) {
      e.p[0] = mkpt(0.0, p0.y);
      e.p[1] = mid;
      e.p[2] = p1;
    } else {
      e.p[0] = p0;
      e.p[1] = mid;
      e.p[2] = mkpt(0.0, p1.y);
    }
    e.n = 2;
  }
  return e;
}

// ActiveEdgeRasterizer's left-edge quirk (src/rasterize.rs:733-737, 760-871,
// 923-937). ActiveEdgeIter walks each edge row by row from its lower end,
// computing the x where it leaves each row incrementally (x += dxdy * dy). A
// row piece whose leftmost x comes out negative (a left-split point at
// x = -1e-17, or the walk arriving at x = 0 a rounding below it) starts its
// column walk with next_ceil(x) = 0, so all of that row piece's deposits
// land one column right of their exact place. Those deposits are the exact-
// area deposits of the row piece moved by x + 1, so an edge line is replaced
// by an equivalent chain: the rows of each piece of its left split as runs,
// those with a negative leftmost x moved by + 1, joined by horizontal lines
// (which deposit nothing). Only pieces whose x comes within kAeNear of 0 are
// walked (elsewhere the incremental x stays far from 0 for any image below
// about 10^5 px); `line` receives the chain's lines in order.
constexpr double kAeNear = 1e-6;
constexpr int kAeRuns = 4;  // runs recorded per piece (more: the rest keeps the last state)
constexpr double kEps = 2.220446049250313e-16;  // f64::EPSILON (src/geometry.rs:14)

// The rows of the piece a -> b (one ActiveEdgeIter edge) below image row H:
// vertices after a up to (not including) b, through `to`.
template <typename To>
RZ_HD void ae_piece(Pt a, Pt b, double H, To&& to) {
  if (fabs(a.y - b.y) < kEps) return;  // Edge::new: no rows
  const bool up = a.y <= b.y;
  const Pt lo = up ? a : b, hi = up ? b : a;
  const double dxdy = (hi.x - lo.x) / (hi.y - lo.y);
  Pt p = lo;
  if (p.y < 0.0) p = mkpt(p.x - p.y * dxdy, 0.0);
  if (!(fmin(p.x, hi.x) < kAeNear)) return;
  Pt rs[kAeRuns];
  bool rf[kAeRuns];
  int nr = 0;
  bool any = false;
  while (hi.y - p.y >= kEps && floor(p.y) < H) {
    const double ys = fmin(floor(p.y) + 1.0, hi.y);
    const Pt q = mkpt(p.x + dxdy * (ys - p.y), ys);
    const bool f = fmin(p.x, q.x) < 0.0;
    if ((nr == 0 || f != rf[nr - 1]) && nr < kAeRuns) {
      rs[nr] = p;
      rf[nr] = f;
      ++nr;
    }
    any = any || f;
    p = q;
  }
  if (!any) return;
  const Pt e = p;  // where the walk ended
  auto sh = [](Pt v) { return mkpt(v.x + 1.0, v.y); };
  if (up) {
    bool st = false;
    for (int k = 0; k < nr; ++k) {
      if (rf[k] == st) continue;
      if (rf[k]) {
        to(rs[k]);
        to(sh(rs[k]));
      } else {
        to(sh(rs[k]));
        to(rs[k]);
      }
      st = rf[k];
    }
    if (st) {
      to(sh(e));
      to(e);
    }
  } else {
    bool st = false;
    if (rf[nr - 1]) {
      to(e);
      to(sh(e));
      st = true;
    }
    for (int k = nr - 1; k >= 0; --k) {
      const bool below = k > 0 ? rf[k - 1] : false;
      if (rf[k] == below) continue;
      if (rf[k]) {
        to(sh(rs[k]));
        to(rs[k]);
      } else {
        to(rs[k]);
        to(sh(rs[k]));
      }
      st = below;
    }
    (void)st;
  }
}

// The chain that replaces the line p0 -> p1 under ActiveEdgeRasterizer
// (H: image rows); `line(a, b)` receives its lines.
template <typename LineF>
RZ_HD void ae_chain(Pt p0, Pt p1, double H, LineF&& line) {
  if (!(fmin(p0.x, p1.x) < kAeNear) || (p0.x <= 0.0 && p1.x <= 0.0)) {
    line(p0, p1);  // never near x = 0, or all projected onto x = 0 exactly
    return;
  }
  Pt prev = p0;
  auto to = [&](Pt v) {
    if (v.x == prev.x && v.y == prev.y) return;
    line(prev, v);
    prev = v;
  };
  if (p0.x >= 0.0 && p1.x >= 0.0) {
    ae_piece(p0, p1, H, to);
    to(p1);
    return;
  }
  // the left split, as effective_line computes it (src/rasterize.rs:923-937)
  const double t = p0.x / (p0.x - p1.x);
  const Pt mid = mkpt((1.0 - t) * p0.x + t * p1.x, (1.0 - t) * p0.y + t * p1.y);
  if (p0.x < 0.0) {
    const Pt v0 = mkpt(0.0, p0.y);
    to(v0);
    ae_piece(v0, mid, H, to);
    to(mid);
    ae_piece(mid, p1, H, to);
    to(p1);
  } else {
    const Pt v1 = mkpt(0.0, p1.y);
    ae_piece(p0, mid, H, to);
    to(mid);
    ae_piece(mid, v1, H, to);
    to(v1);
    to(p1);
  }
}

RZ_HD int tile_col(double x, const BinGrid& g) {
  double f = floor((x + (double)g.phx) / (double)kTile);
  if (!(f >= 0.0)) return 0;
  if (f > (double)(g.tiles_x - 1)) return g.tiles_x - 1;
  return (int)f;
}

// One tile record: a piece of a line clipped to a 16x16 cell, in coordinates
// relative to the cell's top-left corner. The sign bit of x0 (x1) marks that
// the piece starts (ends) where its band piece leaves (enters) the region left
// of the cell (x < c0; x <= 0 for tile 0); the fine stage then adds +V(y0)
// (-V(y1)) to the cell's rows, V(y) being the overlap of each pixel row with
// [0, y].
struct Rec {
  float x0, y0, x1, y1;
};

// Free-endpoint event: a chain endpoint that is not shared with another line
// (a gap between segments, or a vertex created by the right clip). Adds
// sign * V(yrel) to every tile of band `band` right of column `col` (col = -1:
// every tile of the band).
struct Event {
  int band;
  int col;
  float yrel;
  float sign;
};

// The cell decomposition of one effective piece a -> b in three steps that
// can run on different threads: the piece's band range, then per band its
// part inside the band (and the band-bottom crossing), then per tile column
// of that part its record. bin_piece composes them sequentially.

// Bands whose closed y-range meets [ylo, yhi] of the piece: [*ty0, *ty1],
// possibly empty. Returns false when the piece is entirely above or below the
// image.
RZ_HD bool piece_bands(Pt a, Pt b, const BinGrid& g, int* ty0, int* ty1) {
  const double H = (double)g.H;
  const double ylo = fmin(a.y, b.y), yhi = fmax(a.y, b.y);
  if (!(yhi >= 0.0) || !(ylo <= H)) return false;
  int t0 = ylo <= 0.0 ? 0 : (int)ceil((ylo + (double)g.phy) / (double)kTile) - 1;
  if (t0 < 0) t0 = 0;
  int t1 = (int)floor((fmin(yhi, H) + (double)g.phy) / (double)kTile);
  if (t1 > g.tiles_y - 1) t1 = g.tiles_y - 1;
  *ty0 = t0;
  *ty1 = t1;
  return true;
}

// The piece within band ty: its crossing of the band's bottom edge (column of
// the first tile whose left region holds it, -1 if none or beyond the last
// tile) and its part s -> e inside r0 <= y <= rb with that part's tile
// columns [tx0, tx1] (any = false when the part is empty or a single point).
struct BandPiece {
  Pt s, e;
  double r0;
  int tx0, tx1;
  int cross_col;
  bool any;
};

RZ_HD void band_piece(Pt a, Pt b, const BinGrid& g, int ty, BandPiece* bp) {
  const double H = (double)g.H;
  const double ylo = fmin(a.y, b.y), yhi = fmax(a.y, b.y);
  const double r0 = (double)(ty * kTile - g.phy);
  const double rb = fmin(r0 + (double)kTile, H);
  bp->r0 = r0;
  bp->cross_col = -1;
  bp->any = false;
  // band-bottom crossing (a point on y = rb counts as inside the band)
  if (ylo <= rb && rb < yhi) {
    const Pt c = at_y(a, b, rb);
    // first tile whose left region holds c (tile 0's is x <= 0, see tile_record)
    const int col = c.x <= 0.0 ? 0 : tile_col(c.x, g) + 1;
    if (col < g.tiles_x) bp->cross_col = col;
  }
  // piece of a -> b inside r0 <= y <= rb, direction preserved
  Pt s = a, e = b;
  if (a.y != b.y) {
    if (a.y < b.y) {
      if (s.y < r0) s = at_y(a, b, r0);
      if (e.y > rb) e = at_y(a, b, rb);
    } else {
      if (s.y > rb) s = at_y(a, b, rb);
      if (e.y < r0) e = at_y(a, b, r0);
    }
  } else if (a.y < r0 || a.y > rb) {
    return;
  }
  if (s.x == e.x && s.y == e.y) return;  // touches the band in one point only
  bp->s = s;
  bp->e = e;
  bp->tx0 = tile_col(fmin(s.x, e.x), g);
  bp->tx1 = tile_col(fmax(s.x, e.x), g);
  bp->any = true;
}

// The record of band part s -> e (band top r0) in tile column tx, if any.
RZ_HD bool tile_record(Pt s, Pt e, double r0, int tx, const BinGrid& g, Rec* out) {
  const double xlo = fmin(s.x, e.x), xhi = fmax(s.x, e.x);
  const double c0 = (double)(tx * kTile - g.phx), c1 = c0 + (double)kTile;
  if (xhi < c0 || xlo > c1) return false;
  // Tile 0's left region is x <= 0 rather than x < 0: a line on x = 0
  // deposits its whole cover into pixel 0, exactly like a line left of the
  // image, so the vertical lines made by the left split (every part of the
  // path left of the image) need no records at all.
  const bool first = tx == 0;
  if (first && xhi <= 0.0) return false;
  // clip the band part to c0 <= x <= c1
  Pt u = s, v = e;
  if (s.x != e.x) {
    if (s.x < e.x) {
      if (u.x < c0) u = at_x(s, e, c0);
      if (v.x > c1) v = at_x(s, e, c1);
    } else {
      if (u.x > c1) u = at_x(s, e, c1);
      if (v.x < c0) v = at_x(s, e, c0);
    }
  }
  // crossings of the cell's left edge x = c0 (inside means x < c0)
  const bool in_s = first ? s.x <= 0.0 : s.x < c0;
  const bool in_e = first ? e.x <= 0.0 : e.x < c0;
  bool flag0 = false, flag1 = false;
  if (in_s && !in_e) flag0 = u.y != r0;  // exits the left region: +V(u.y)
  if (!in_s && in_e) flag1 = v.y != r0;  // enters the left region: -V(v.y)
  // no vertical extent and no edge crossing: nothing to deposit
  if (u.y == v.y && !flag0 && !flag1) return false;
  Rec r;
  r.x0 = (float)(u.x - c0);
  r.y0 = (float)(u.y - r0);
  r.x1 = (float)(v.x - c0);
  r.y1 = (float)(v.y - r0);
  // the flag is the sign bit; the magnitude stays the piece's x, which is 0
  // except where tile 0's left region (x <= 0) ends inside the tile (phx > 0)
  if (flag0) r.x0 = -r.x0;
  if (flag1) r.x1 = -r.x1;
  *out = r;
  return true;
}

// Enumerate the records and band-boundary crossings of one effective piece
// a -> b, band by band and left to right.
//   rec(tile_index, Rec)        -- one cell record
//   bd(band, col, dir)          -- crossing of the band's bottom edge at column col
template <typename RecF, typename BdF>
RZ_HD void bin_piece(Pt a, Pt b, const BinGrid& g, RecF&& rec, BdF&& bd) {
  int ty0, ty1;
  if (!piece_bands(a, b, g, &ty0, &ty1)) return;
  const int dir = b.y > a.y ? 1 : -1;
  for (int ty = ty0; ty <= ty1; ++ty) {
    BandPiece bp;
    band_piece(a, b, g, ty, &bp);
    if (bp.cross_col >= 0) bd(ty, bp.cross_col, dir);
    if (!bp.any) continue;
    for (int tx = bp.tx0; tx <= bp.tx1; ++tx) {
      Rec r;
      if (tile_record(bp.s, bp.e, bp.r0, tx, g, &r)) rec(ty * g.tiles_x + tx, r);
    }
  }
}

// Right-edge column. A line beyond the right clip becomes a vertical line on
// X = redge_x(g) (src/rasterize.rs:372-375). Per pixel row it deposits
// d (1 - f) into pixel floor(X) and d f into the next one (f = X - floor(X),
// src/rasterize.rs:437-444), d being the row's signed overlap V(y1) - V(y0),
// V(y)(row) = clamp(y - row, 0, 1). All such lines share X, so only the per-row
// sum R(row) of d matters: pixel floor(X) gets (1 - f) R and every pixel right
// of it R. Such lines therefore make no records, crossings or events; each adds
// its two endpoint terms to E, the 2^-20 fixed-point differences of V between
// consecutive rows (integer sums, exact in any order), and the row prefix of E
// is R.
RZ_HD int fix_frac(double f) { return (int)rint(f * (double)(1 << kFixShift)); }

// sign * V(y) as differences E[k] += sign (F - 1), E[k + 1] -= sign F with
// k = floor(y), F = frac(y), y clamped to [0, H] (E has H + 2 entries; the
// constant term of V cancels between a line's two ends).
template <typename AddF>
RZ_HD void redge_endpoint(double y, int sign, const BinGrid& g, AddF&& add) {
  const double H = (double)g.H;
  const double yc = y < 0.0 ? 0.0 : (y > H ? H : y);
  const int k = (int)floor(yc);
  const int F = fix_frac(yc - (double)k);
  const int one = 1 << kFixShift;
  if (F != one) add(k, sign * (F - one));
  if (F != 0) add(k + 1, -sign * F);
}

// Free-endpoint event for effective vertex p (sign +1 for a chain end, -1 for
// a chain start). Returns false when no tile lies right of it in its band or
// the point is outside the image rows (then the event has no effect).
RZ_HD bool make_event(Pt p, float sign, const BinGrid& g, Event* ev) {
  const double H = (double)g.H;
  if (!(p.y > 0.0) || !(p.y <= H)) return false;
  int band = (int)ceil((p.y + (double)g.phy) / (double)kTile) - 1;
  if (band < 0) band = 0;
  if (band > g.tiles_y - 1) return false;
  // an endpoint on x = 0 lies in tile 0's left region: it affects every tile
  int col = p.x <= 0.0 ? -1 : tile_col(p.x, g);
  if (col >= g.tiles_x - 1) return false;
  ev->band = band;
  ev->col = col;
  ev->yrel = (float)(p.y - (double)(band * kTile - g.phy));
  ev->sign = sign;
  return true;
}

}  // namespace rzg
