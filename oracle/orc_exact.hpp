// Oracle: FP64 exact-area checkers, independent of the reference's formulas.
//
// TEST INFRASTRUCTURE ONLY (see orc_core.hpp).
//
//  * exact_winding_area: per pixel, the winding-weighted area of closed rings
//    intersected with the pixel square, by Sutherland-Hodgman clipping against
//    the four pixel half-planes followed by the shoelace formula. Clipping a
//    ring against a half-plane replaces every outside chain by a segment on the
//    boundary line, which leaves the winding number of every inside point
//    unchanged, so the shoelace area of the clipped ring is exactly
//    integral(pixel) w dA. This equals the reference's accumulated signed
//    difference before the fill rule (box-filtered winding), in real arithmetic.
//
//  * exact_fill_coverage: per pixel, the exact area of {p in pixel : rule(w(p))}
//    for a set of lines, by slab decomposition: y is broken at every vertex,
//    every pairwise edge intersection and every integer; within a slab no two
//    edges cross, so edges are ordered by x, the winding number between
//    neighbours is constant, and each inside region is a trapezoid whose
//    intersection with a pixel column integrates exactly (piecewise linear).
//    This is the coverage of the filled region itself, i.e. what an ideal
//    anti-aliased rasterizer would produce; the reference (and the HIP path)
//    apply the fill rule after box filtering, which differs where edges of
//    different subpaths overlap inside one pixel.
#pragma once
#include <algorithm>
#include "orc_core.hpp"

namespace orc {

// Clip polygon `in` against the half-plane {sign*(coord - c) >= 0} on axis
// (0 = x, 1 = y).
inline void sh_clip(const std::vector<Point>& in, int axis, double c, double sign,
                    std::vector<Point>& out) {
  out.clear();
  size_t n = in.size();
  if (n == 0) return;
  auto val = [&](const Point& p) { return sign * ((axis == 0 ? p.x : p.y) - c); };
  for (size_t i = 0; i < n; ++i) {
    const Point& a = in[i];
    const Point& b = in[(i + 1) % n];
    double va = val(a), vb = val(b);
    bool ia = va >= 0.0, ib = vb >= 0.0;
    if (ia) out.push_back(a);
    if (ia != ib) {
      double t = va / (va - vb);
      Point q = P(a.x + t * (b.x - a.x), a.y + t * (b.y - a.y));
      if (axis == 0) q.x = c; else q.y = c;
      out.push_back(q);
    }
  }
}

inline double shoelace(const std::vector<Point>& p) {
  double s = 0.0;
  size_t n = p.size();
  for (size_t i = 0; i < n; ++i) {
    const Point& a = p[i];
    const Point& b = p[(i + 1) % n];
    s += a.x * b.y - b.x * a.y;
  }
  return 0.5 * s;
}

// rings: closed polylines (last point connects to the first). Output: W*H
// row-major accumulated winding (positive for rings that run clockwise in
// image coordinates, matching the reference's sign: a downward edge adds
// coverage to its right).
inline void exact_winding_area(const std::vector<std::vector<Point>>& rings, size_t W, size_t H,
                               double* out) {
  std::vector<Point> band, a, b;
  for (size_t y = 0; y < H; ++y) {
    for (size_t x = 0; x < W; ++x) out[y * W + x] = 0.0;
    for (const auto& ring : rings) {
      sh_clip(ring, 1, (double)y, 1.0, a);
      sh_clip(a, 1, (double)(y + 1), -1.0, band);
      if (band.size() < 3) continue;
      double bxmin = band[0].x, bxmax = band[0].x;
      for (const Point& p : band) { bxmin = std::min(bxmin, p.x); bxmax = std::max(bxmax, p.x); }
      long x0 = std::max(0L, (long)std::floor(bxmin));
      long x1 = std::min((long)W - 1, (long)std::floor(bxmax));
      for (long x = x0; x <= x1; ++x) {
        sh_clip(band, 0, (double)x, 1.0, a);
        sh_clip(a, 0, (double)(x + 1), -1.0, b);
        if (b.size() < 3) continue;
        out[y * W + (size_t)x] += shoelace(b);
      }
      // pixels left of the ring's band extent are outside; pixels right of it
      // are outside as well (a closed ring has winding 0 beyond its extent).
    }
  }
  // Shoelace is positive for counter-clockwise rings in a y-up frame; in the
  // image frame (y down) a ring whose downward edges are on the left, i.e.
  // which the reference fills with +1, has negative shoelace area.
  for (size_t i = 0; i < W * H; ++i) out[i] = -out[i];
}

// integral over y in [ya, yb] of clamp(x(y) - p, 0, 1) with x linear in y,
// x(ya) = xa, x(yb) = xb. Exact: splits at the breakpoints x = p, x = p + 1.
inline double int_clamp_lin(double ya, double yb, double xa, double xb, double p) {
  double h = yb - ya;
  if (h <= 0.0) return 0.0;
  // parametrize u in [0,1]
  double ts[4] = {0.0, 1.0, 0.0, 0.0};
  int n = 2;
  double dx = xb - xa;
  if (dx != 0.0) {
    double u0 = (p - xa) / dx, u1 = (p + 1.0 - xa) / dx;
    if (u0 > 0.0 && u0 < 1.0) ts[n++] = u0;
    if (u1 > 0.0 && u1 < 1.0) ts[n++] = u1;
  }
  std::sort(ts, ts + n);
  double acc = 0.0;
  for (int i = 0; i + 1 < n; ++i) {
    double ua = ts[i], ub = ts[i + 1];
    double fa = xa + ua * dx - p, fb = xa + ub * dx - p;
    // between breakpoints clamp() is linear, so the trapezoid rule is exact
    double ca = std::min(1.0, std::max(0.0, fa)), cb = std::min(1.0, std::max(0.0, fb));
    acc += 0.5 * (ca + cb) * (ub - ua);
  }
  return acc * h;
}

struct XEdge {
  double xa, xb;  // x at slab top / bottom
  int dir;        // +1 when the edge runs downward (y increasing)
};

// rule: 0 nonzero, 1 evenodd
inline void exact_fill_coverage(const std::vector<Line>& lines, size_t W, size_t H, int rule,
                                double* out) {
  for (size_t i = 0; i < W * H; ++i) out[i] = 0.0;
  std::vector<Line> ls;
  for (const Line& l : lines)
    if (l.p0.y != l.p1.y) ls.push_back(l);
  // breakpoints in y
  std::vector<double> ys;
  for (size_t y = 0; y <= H; ++y) ys.push_back((double)y);
  for (const Line& l : ls) { ys.push_back(l.p0.y); ys.push_back(l.p1.y); }
  for (size_t i = 0; i < ls.size(); ++i) {
    for (size_t j = i + 1; j < ls.size(); ++j) {
      const Line& a = ls[i];
      const Line& b = ls[j];
      double x1 = a.p0.x, y1 = a.p0.y, x2 = a.p1.x, y2 = a.p1.y;
      double x3 = b.p0.x, y3 = b.p0.y, x4 = b.p1.x, y4 = b.p1.y;
      double det = (x4 - x3) * (y1 - y2) - (x1 - x2) * (y4 - y3);
      if (det == 0.0) continue;
      double t0 = ((y3 - y4) * (x1 - x3) + (x4 - x3) * (y1 - y3)) / det;
      double t1 = ((y1 - y2) * (x1 - x3) + (x2 - x1) * (y1 - y3)) / det;
      if (t0 > 0.0 && t0 < 1.0 && t1 > 0.0 && t1 < 1.0) ys.push_back(y1 + t0 * (y2 - y1));
    }
  }
  std::sort(ys.begin(), ys.end());
  ys.erase(std::unique(ys.begin(), ys.end()), ys.end());
  std::vector<XEdge> es;
  for (size_t k = 0; k + 1 < ys.size(); ++k) {
    double ya = ys[k], yb = ys[k + 1];
    if (yb <= 0.0 || ya >= (double)H) continue;
    if (yb - ya <= 0.0) continue;
    size_t row = (size_t)std::floor(0.5 * (ya + yb));
    if (row >= H) continue;
    es.clear();
    double ym = 0.5 * (ya + yb);
    for (const Line& l : ls) {
      double lo = std::min(l.p0.y, l.p1.y), hi = std::max(l.p0.y, l.p1.y);
      if (!(lo < ym && ym < hi)) continue;
      double dxdy = (l.p1.x - l.p0.x) / (l.p1.y - l.p0.y);
      XEdge e;
      e.xa = l.p0.x + (ya - l.p0.y) * dxdy;
      e.xb = l.p0.x + (yb - l.p0.y) * dxdy;
      e.dir = l.p1.y > l.p0.y ? 1 : -1;
      es.push_back(e);
    }
    std::sort(es.begin(), es.end(), [](const XEdge& a, const XEdge& b) {
      return (a.xa + a.xb) < (b.xa + b.xb);
    });
    int w = 0;
    for (size_t i = 0; i + 1 < es.size(); ++i) {
      w += es[i].dir;
      bool inside = rule == 0 ? (w != 0) : ((w & 1) != 0);
      if (!inside) continue;
      const XEdge& L = es[i];
      const XEdge& R = es[i + 1];
      double xmin = std::min(L.xa, L.xb), xmax = std::max(R.xa, R.xb);
      long c0 = std::max(0L, (long)std::floor(xmin));
      long c1 = std::min((long)W - 1, (long)std::floor(xmax));
      for (long c = c0; c <= c1; ++c) {
        double v = int_clamp_lin(ya, yb, R.xa, R.xb, (double)c) -
                   int_clamp_lin(ya, yb, L.xa, L.xb, (double)c);
        out[row * W + (size_t)c] += v;
      }
    }
  }
}

}  // namespace orc
