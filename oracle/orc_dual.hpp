// Oracle: forward-mode derivative of SignedDifferenceRasterizer::mask_iter
// with respect to one direction in a path's control points.
//
// TEST INFRASTRUCTURE ONLY (see orc_core.hpp). This is the CPU FP64
// reference gradient of the differentiable render: the f64 algorithm of
// orc_core.hpp / orc_raster.hpp (src/path.rs:734-796, src/rasterize.rs:313-
// 470) run on (value, derivative) pairs. Every decision -- subdivision,
// clipping, pixel rows and columns, the fill rule's branch -- is taken on the
// values, exactly as the plain oracle takes it, so the values reproduce
// sd_mask_iter bit for bit and the derivatives are those of the same
// piecewise-smooth function at fixed decisions. It shares no code with the
// product's adjoint (which integrates the winding's gradient along the edges).
#pragma once
#include "orc_raster.hpp"

namespace orc {

struct Dual {
  double v, d;
};
inline Dual D(double v, double d = 0.0) { return Dual{v, d}; }
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
inline Dual dmax(Dual a, Dual b) { return a.v >= b.v ? a : b; }  // std::fmax on values
inline Dual dmin(Dual a, Dual b) { return a.v <= b.v ? a : b; }
inline Dual dfloor(Dual a) { return {std::floor(a.v), 0.0}; }
inline Dual dceil(Dual a) { return {std::ceil(a.v), 0.0}; }

struct DPoint {
  Dual x, y;
};
struct DLine {
  DPoint p0, p1;
};

// signed_difference_line (src/rasterize.rs:365-470) on dual numbers, for the
// dense (W+1)-wide scratch of mask_iter: add(offset, value, derivative).
template <typename Add>
inline void sd_line_deposits_dual(const Shape& shape, size_t len, DLine line, Add&& add) {
  DPoint p0 = line.p0, p1 = line.p1;
  const Dual width = D((double)shape.width - 1.0);
  if (p0.x.v > width.v || p1.x.v > width.v) {
    if (p0.x.v > width.v && p1.x.v > width.v) {
      line = DLine{DPoint{width - D(0.001), p0.y}, DPoint{width - D(0.001), p1.y}};
    } else {
      Dual t = (p0.x - width) / (p0.x - p1.x);
      DPoint mid{width, (D(1.0) - t) * p0.y + t * p1.y};
      if (p0.x.v < width.v) line = DLine{p0, mid};
      else line = DLine{mid, p1};
    }
  }
  // split_at_zero_x (src/rasterize.rs:922-937)
  {
    DPoint a = line.p0, b = line.p1;
    if (!(a.x.v >= 0.0 && b.x.v >= 0.0)) {
      if (a.x.v <= 0.0 && b.x.v <= 0.0) {
        line = DLine{DPoint{D(0.0), a.y}, DPoint{D(0.0), b.y}};
      } else {
        Dual t = a.x / (a.x - b.x);
        DPoint mid{(D(1.0) - t) * a.x + t * b.x, (D(1.0) - t) * a.y + t * b.y};
        DLine rest;
        if (a.x.v < 0.0) {
          line = DLine{mid, b};
          rest = DLine{DPoint{D(0.0), a.y}, mid};
        } else {
          line = DLine{a, mid};
          rest = DLine{mid, DPoint{D(0.0), b.y}};
        }
        sd_line_deposits_dual(shape, len, rest, add);
      }
    }
  }
  p0 = line.p0;
  p1 = line.p1;
  const size_t stride = shape.col_stride;
  if (std::fabs(p0.y.v - p1.y.v) < EPSILON) return;
  Dual dir;
  if (p0.y.v < p1.y.v) {
    dir = D(1.0);
  } else {
    dir = D(-1.0);
    std::swap(p0, p1);
  }
  Dual dxdy = (p1.x - p0.x) / (p1.y - p0.y);
  size_t y = (size_t)std::fmax(p0.y.v, 0.0);
  Dual x = p0.y.v < 0.0 ? p0.x - p0.y * dxdy : p0.x;
  Dual x_next = x;
  double yend_f = std::fmax(std::ceil(p1.y.v), 0.0);
  size_t yend = std::min(shape.height, (size_t)yend_f);
  for (; y < yend; ++y) {
    x = x_next;
    size_t row_offset = shape.offset(y, 0);
    Dual dy = dmin(D((double)(y + 1)), p1.y) - dmax(D((double)y), p0.y);
    Dual d = dir * dy;
    x_next = x + dxdy * dy;
    Dual x0, x1;
    if (x.v < x_next.v) { x0 = x; x1 = x_next; } else { x0 = x_next; x1 = x; }
    Dual x0_floor = dmax(dfloor(x0), D(0.0));
    int x0i = (int)x0_floor.v;
    Dual x1_ceil = dmin(dceil(x1), width);
    int x1i = (int)x1_ceil.v;
    auto put = [&](size_t off, Dual v) { add(off, v.v, v.d); };
    if (x1i <= x0i + 1) {
      Dual xmf = D(0.5) * (x + x_next) - x0_floor;
      put(row_offset + (size_t)x0i * stride, d * (D(1.0) - xmf));
      size_t offset = row_offset + (size_t)(x0i + 1) * stride;
      if (offset < len) put(offset, d * xmf);
    } else {
      Dual s = D(1.0) / (x1 - x0);
      Dual x0f = x0 - x0_floor;
      Dual x1f = x1 - x1_ceil + D(1.0);
      Dual a0 = D(0.5) * s * (D(1.0) - x0f) * (D(1.0) - x0f);
      Dual am = D(0.5) * s * x1f * x1f;
      put(row_offset + (size_t)x0i * stride, d * a0);
      if (x1i == x0i + 2) {
        put(row_offset + (size_t)(x0i + 1) * stride, d * (D(1.0) - a0 - am));
      } else {
        Dual a1 = s * (D(1.5) - x0f);
        put(row_offset + (size_t)(x0i + 1) * stride, d * (a1 - a0));
        for (int xi = x0i + 2; xi < x1i - 1; ++xi) put(row_offset + (size_t)xi * stride, d * s);
        Dual a2 = a1 + D((double)(x1i - x0i - 3)) * s;
        put(row_offset + (size_t)(x1i - 1) * stride, d * (D(1.0) - a2 - am));
      }
      put(row_offset + (size_t)x1i * stride, d * am);
    }
  }
}

// Path::flatten (src/path.rs:734-796) carrying a tangent: `tan` holds, for
// every stored point of the path (point index order), the derivative of that
// point's x and y. Subdivision is decided on the values; split_at(0.5) and
// the transform are linear, so the tangents follow the same maps.
template <typename Emit>
inline void flatten_dual(const Path& path, const double* tan, const Transform& tr, double flatness,
                         Emit&& emit) {
  const double fl = 16.0 * flatness * flatness;
  auto lin = [&](Point t) {  // the transform's linear part
    return P(t.x * tr.m[0] + t.y * tr.m[1], t.x * tr.m[3] + t.y * tr.m[4]);
  };
  std::vector<std::pair<Segment, Segment>> stack;
  size_t pt = 0;
  std::vector<size_t> first(path.segments.size());
  for (size_t k = 0; k < path.segments.size(); ++k) {
    first[k] = pt;
    pt += (size_t)path.segments[k].npts();
  }
  auto dline = [](Point a, Point da, Point b, Point db) {
    return DLine{DPoint{D(a.x, da.x), D(a.y, da.y)}, DPoint{D(b.x, db.x), D(b.y, db.y)}};
  };
  for (size_t i = 0; i < path.nsubpaths(); ++i) {
    size_t b = path.subpaths[i], e = path.subpaths[i + 1];
    for (size_t k = b; k < e; ++k) {
      const Segment& s0 = path.segments[k];
      Segment t0 = s0;
      for (int j = 0; j < s0.npts(); ++j)
        t0.p[j] = lin(P(tan[2 * (first[k] + j)], tan[2 * (first[k] + j) + 1]));
      stack.push_back({s0.transform(tr), t0});
      while (!stack.empty()) {
        auto st = stack.back();
        stack.pop_back();
        const Segment& s = st.first;
        if (s.has_nans()) throw std::runtime_error("cannot flatten segment with NaN");
        if (s.flatness() < fl) {
          emit(dline(s.start(), st.second.start(), s.end(), st.second.end()));
          continue;
        }
        Segment a, c, ta, tc;
        s.split(&a, &c);
        st.second.split(&ta, &tc);
        stack.push_back({c, tc});
        stack.push_back({a, ta});
      }
    }
    // the closing line, as Path::flatten with close = true
    const Segment& se = path.segments[e - 1];
    const Segment& sb = path.segments[b];
    const size_t ie = first[e - 1] + (size_t)se.npts() - 1, ib = first[b];
    emit(dline(tr.apply(se.end()), lin(P(tan[2 * ie], tan[2 * ie + 1])), tr.apply(sb.start()),
               lin(P(tan[2 * ib], tan[2 * ib + 1]))));
  }
}

// SignedDifferenceRasterizer::mask_iter (src/rasterize.rs:313-355) and its
// derivative: alpha[y * W + x] and dalpha[y * W + x] for the pixels it
// yields (|alpha| >= 1e-6), zero elsewhere. FillRule::alpha_from_winding
// (src/path.rs:32-46) is differentiated on the branch the value takes.
inline void sd_mask_iter_dual(const Path& path, const double* tan, const Transform& tr,
                              double flatness, FillRule rule, size_t width, size_t height,
                              double* alpha, double* dalpha) {
  std::fill(alpha, alpha + width * height, 0.0);
  std::fill(dalpha, dalpha + width * height, 0.0);
  if (width == 0 || height == 0) return;
  const size_t w1 = width + 1;
  std::vector<double> sv(w1 * height, 0.0), sd(w1 * height, 0.0);
  const Shape shape = Shape::simple(height, w1);
  flatten_dual(path, tan, tr, flatness, [&](const DLine& l) {
    sd_line_deposits_dual(shape, sv.size(), l, [&](size_t off, double v, double d) {
      sv[off] += v;
      sd[off] += d;
    });
  });
  double w = 0.0, dw = 0.0;
  for (size_t index = 0; index < sv.size(); ++index) {
    size_t y = index / w1;
    size_t x = index - y * w1;
    if (x == 0) {
      w = 0.0;
      dw = 0.0;
    } else if (x >= width) {
      continue;
    }
    w += sv[index];
    dw += sd[index];
    double a = alpha_from_winding(rule, w);
    if (std::fabs(a) < 1e-6) continue;
    double da;
    if (rule == EVENODD) {
      // |rem_euclid(w + 1, 2) - 1|: slope +-1
      double r = rem_euclid(w + 1.0, 2.0) - 1.0;
      da = r >= 0.0 ? dw : -dw;
    } else {
      da = std::fabs(w) >= 1.0 ? 0.0 : (w >= 0.0 ? dw : -dw);
    }
    alpha[y * width + x] = a;
    dalpha[y * width + x] = da;
  }
}

}  // namespace orc
